// Kernel (b): the right-upper triangular solve Y <- Y R^-1 that replaces
// rdiv_upper! (src/kernels.jl:108, semantics :90-99, called from rdiv_rows!
// src/panelblas.jl:157-169), and its conjugate-transpose twin Y <- Y R^-H
// for the QR adjoint.
//
// Design (gfx950). Rows are independent, so a workgroup owns TW_ROWS rows and
// walks the columns in blocks of TW_CB, left looking:
//     X_J = (Y_J - sum_{I<J} X_I R_IJ) * inv(R_JJ).
// Wide blocks keep HBM traffic down: the update re-reads the workgroup's
// solved columns once per block, so per row it reads sum_J TW_CB J columns,
// plus k for the right-hand side and k for the diagonal multiply (at k = 512:
// 2.8 k for 64-wide blocks, 8.5 k for 32-wide ones). The inverses of the
// diagonal blocks are formed once per call, in double, by
// trsm_block_inv_kernel (a right-looking sweep over the rows of the inverse,
// one barrier per row). The update and the multiply by the
// diagonal inverse both run as three-product (3M) complex matrix-core steps
// over LDS-staged 16-column slices, the next slice's loads in flight during
// the current slice's products. Between the two, the block's right-hand side
// makes one round trip through the workgroup's own rows of Y, which are then
// hot in L2.
#include "fhip_internal.hpp"

#include <type_traits>

namespace fhip {

// Column-block width; a workgroup owns 4096 / TW_CB rows, so that its four
// waves each hold a 32 x 32 tile of the rows x TW_CB block.
#ifndef FHIP_TRSM_CB
#define FHIP_TRSM_CB 64
#endif
constexpr int TW_THREADS = 256;  // 4 waves
constexpr int TW_CB = FHIP_TRSM_CB;
constexpr int TW_ROWS = 4096 / TW_CB;
// Inner dimension per LDS slice (a tuning knob per compute precision; 32
// for FP32 measured slower than 16 at 629856 x 512).
#ifndef FHIP_TRSM_KC_FP32
#define FHIP_TRSM_KC_FP32 16
#endif
template <typename T>
struct TrsmSlice {
  static constexpr int KC = sizeof(T) == 4 ? FHIP_TRSM_KC_FP32 : 16;
  static constexpr int APT = KC * TW_ROWS / TW_THREADS;  // A elements per thread per slice
  static constexpr int BPT = KC * TW_CB / TW_THREADS;    // B elements per thread per slice
};
constexpr int TW_LDA = TW_ROWS + 4;
constexpr int TW_LDB = TW_CB + 4;
constexpr int TI_THREADS = 8 * TW_CB;                 // block inverse: 8 row groups x TW_CB columns
constexpr int TI_RPT = TW_CB / 8;                     // rows of the inverse per thread
constexpr int TI_CHUNK = 16;                          // columns of R staged in LDS at a time
static_assert(TW_CB == 32 || TW_CB == 64 || TW_CB == 128, "column block of 32, 64 or 128");

// Inverse of each TW_CB x TW_CB diagonal block of R (the last one cut to
// what is left of k, zero padded), column major, one block after another,
// computed in double. W = inv(R_JJ) solves R_JJ W = I from the bottom row
// up: once row l of W is known, its multiple R[i, l] W[l, :] is taken off
// every row i < l. Thread (g, j) keeps column j of rows g, g + 8, ... in
// registers; R comes through LDS sixteen columns at a time.
template <typename TS, typename T, bool CPLX>
__global__ __launch_bounds__(TI_THREADS) void trsm_block_inv_kernel(const TS *__restrict__ R, int64_t ldr, int k,
                                                                     T *__restrict__ Dinv) {
  constexpr int NP = CPLX ? 2 : 1;
  __shared__ double rc_re[TI_CHUNK][TW_CB], rc_im[TI_CHUNK][TW_CB];
  __shared__ double row_re[2][TW_CB], row_im[2][TW_CB];
  const int J = blockIdx.x;
  const int c0 = J * TW_CB;
  const int nb = (k - c0) < TW_CB ? (k - c0) : TW_CB;
  const int t = threadIdx.x;
  const int j = t % TW_CB;
  const int g = t / TW_CB;
  double br[TI_RPT], bi[TI_RPT];
#pragma unroll
  for (int q = 0; q < TI_RPT; ++q) {
    br[q] = (g + 8 * q == j && j < nb) ? 1.0 : 0.0;
    bi[q] = 0.0;
  }
  for (int lc = (nb - 1) / TI_CHUNK; lc >= 0; --lc) {
    const int l0 = lc * TI_CHUNK;
    // Columns [l0, l0 + 16) of the block, rows i <= l (zero elsewhere).
    __syncthreads();
    for (int e = t; e < TI_CHUNK * TW_CB; e += TI_THREADS) {
      const int i = e % TW_CB;
      const int ll = e / TW_CB;
      const int l = l0 + ll;
      double re = 0.0, im = 0.0;
      if (l < nb && i <= l) {
        const int64_t idx = (int64_t)(c0 + i) + (int64_t)(c0 + l) * ldr;
        re = (double)R[CPLX ? 2 * idx : idx];
        im = CPLX ? (double)R[2 * idx + 1] : 0.0;
      }
      rc_re[ll][i] = re;
      rc_im[ll][i] = im;
    }
    __syncthreads();
    const int ltop = (l0 + TI_CHUNK - 1 < nb - 1) ? l0 + TI_CHUNK - 1 : nb - 1;
    for (int l = ltop; l >= l0; --l) {
      const int p = l & 1;
      const int ll = l - l0;
      // Row l of W is final: divide by the pivot and publish it.
      if (g == (l & 7)) {
        const double dr = rc_re[ll][l], di = rc_im[ll][l];
        const double d = dr * dr + di * di;
#pragma unroll
        for (int q = 0; q < TI_RPT; ++q)
          if (q == (l >> 3)) {
            const double wr = (br[q] * dr + bi[q] * di) / d;
            const double wi = (bi[q] * dr - br[q] * di) / d;
            br[q] = wr;
            bi[q] = wi;
            row_re[p][j] = wr;
            row_im[p][j] = wi;
          }
      }
      // The row buffers alternate, so one barrier per row suffices: a row's
      // buffer is rewritten two rows later, after every thread has passed
      // the barrier of the row in between.
      __syncthreads();
      const double wr = row_re[p][j], wi = row_im[p][j];
#pragma unroll
      for (int q = 0; q < TI_RPT; ++q) {
        const int i = g + 8 * q;
        if (i < l) {
          const double cr = rc_re[ll][i], ci = rc_im[ll][i];
          br[q] -= cr * wr - ci * wi;
          bi[q] -= cr * wi + ci * wr;
        }
      }
    }
  }
  T *Z = Dinv + (int64_t)J * TW_CB * TW_CB * NP;
#pragma unroll
  for (int q = 0; q < TI_RPT; ++q) {
    const int i = g + 8 * q;
    const bool keep = i <= j && j < nb;
    Z[(i + (int64_t)j * TW_CB) * NP] = keep ? (T)br[q] : T(0);
    if (CPLX) Z[(i + (int64_t)j * TW_CB) * NP + 1] = keep ? (T)bi[q] : T(0);
  }
}

// The solve. CT = false solves X R = Y (left to right over column blocks);
// CT = true solves X R^H = Y, whose factor R^H is lower triangular, right to
// left. Each wave computes a 32 x 32 piece of the workgroup's
// TW_ROWS x TW_CB block as 2 x 2 matrix-core tiles.
template <typename TS, typename T, bool CPLX, bool CT>
__global__ __launch_bounds__(TW_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void trsm_wide_kernel(
    TS *__restrict__ Y, int64_t ldy, int64_t m, int k, const TS *__restrict__ R, int64_t ldr,
    const T *__restrict__ Dinv) {
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int TW_KC = TrsmSlice<T>::KC;
  constexpr int TW_APT = TrsmSlice<T>::APT;
  constexpr int TW_BPT = TrsmSlice<T>::BPT;
  __shared__ T As[NP][TW_KC * TW_LDA];
  __shared__ T Bs[NP][TW_KC * TW_LDB];
  typedef typename Mfma<T>::acc acc_t;

  const int64_t row0 = (int64_t)blockIdx.x * TW_ROWS;
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  constexpr int WAVE_ROWS = TW_ROWS / 32;  // waves down the rows of the block
  const int mi0 = 32 * (w % WAVE_ROWS);
  const int nj0 = 32 * (w / WAVE_ROWS);
  const int nbk = (k + TW_CB - 1) / TW_CB;

  T ar[TW_APT], ai[TW_APT], bre[TW_BPT], bim[TW_BPT];
  // A slice: columns [kc, kc + 16) of the workgroup's rows of Y (zero at or
  // past column climit).
  auto fetch_a = [&](int kc, int climit) {
#pragma unroll
    for (int e = 0; e < TW_APT; ++e) {
      const int idx = t + TW_THREADS * e;
      const int r = idx % TW_ROWS;
      const int kk = idx / TW_ROWS;
      const int64_t row = row0 + r;
      load_elem<TS, T, CPLX>(Y, row + (int64_t)(kc + kk) * ldy, row < m && kc + kk < climit, ar[e], ai[e]);
    }
  };
  // B slice of the update: rows [kc, kc + 16) x columns [c0, c0 + 128) of R
  // (of R^H when CT), zero past k.
  auto fetch_r = [&](int kc, int c0) {
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const int idx = t + TW_THREADS * e;
      const int kk = idx % TW_KC;
      const int c = idx / TW_KC;
      const bool ok = kc + kk < k && c0 + c < k;
      if (CT) {
        // Element (kc + kk, c0 + c) of R^H is conj(R[c0 + c, kc + kk]).
        load_elem<TS, T, CPLX>(R, (int64_t)(c0 + c) + (int64_t)(kc + kk) * ldr, ok, bre[e], bim[e]);
        bim[e] = -bim[e];
      } else {
        load_elem<TS, T, CPLX>(R, (int64_t)(kc + kk) + (int64_t)(c0 + c) * ldr, ok, bre[e], bim[e]);
      }
    }
  };
  // B slice of the diagonal multiply: rows [kc, kc + 16) of inv(R_JJ) (of
  // inv(R_JJ)^H = inv(R_JJ^H) when CT).
  auto fetch_d = [&](int kc, const T *Dj) {
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const int idx = t + TW_THREADS * e;
      const int kk = idx % TW_KC;
      const int c = idx / TW_KC;
      if (CT) {
        load_elem<T, T, CPLX>(Dj, (int64_t)c + (int64_t)(kc + kk) * TW_CB, true, bre[e], bim[e]);
        bim[e] = -bim[e];
      } else {
        load_elem<T, T, CPLX>(Dj, (int64_t)(kc + kk) + (int64_t)c * TW_CB, true, bre[e], bim[e]);
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int e = 0; e < TW_APT; ++e) {
      const int idx = t + TW_THREADS * e;
      const int r = idx % TW_ROWS;
      const int kk = idx / TW_ROWS;
      As[0][kk * TW_LDA + r] = ar[e];
      if (CPLX) As[NP - 1][kk * TW_LDA + r] = ai[e];
    }
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const int idx = t + TW_THREADS * e;
      const int kk = idx % TW_KC;
      const int c = idx / TW_KC;
      Bs[0][kk * TW_LDB + c] = bre[e];
      if (CPLX) Bs[NP - 1][kk * TW_LDB + c] = bim[e];
    }
  };
  auto mma = [&](acc_t(&p1)[2][2], acc_t(&p2)[2][2], acc_t(&p3)[2][2], auto mode) {
    constexpr int MODE = decltype(mode)::value;
    if constexpr (CPLX)
      mma_steps_3m<T, 2, 2, MODE>(As[0], As[NP - 1], TW_LDA, mi0, Bs[0], Bs[NP - 1], TW_LDB, nj0, 0, TW_KC / 4, 1,
                                  p1, p2, p3);
    else
      mma_steps<T, false, 2, 2, MODE>(As[0], As[0], TW_LDA, mi0, Bs[0], Bs[0], TW_LDB, nj0, 0, TW_KC / 4, 1, p1,
                                      p2);
  };
  typedef std::integral_constant<int, PROD_SUB> Sub;
  typedef std::integral_constant<int, PROD_ADD> Add;

  for (int step = 0; step < nbk; ++step) {
    const int J = CT ? nbk - 1 - step : step;
    const int c0 = J * TW_CB;
    const int cend = (c0 + TW_CB < k) ? c0 + TW_CB : k;
    // Waves whose columns lie past k only stage slices.
    const bool active = c0 + nj0 < k;
    // The block's right-hand side seeds the 3M sums: p1 = Y_re, p2 = 0,
    // p3 = Y_re + Y_im, counted down by the products, so that combine3m
    // yields Y_J - sum X_I R_IJ.
    acc_t p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + nj0 + 16 * nb + li;
          T re, im;
          load_elem<TS, T, CPLX>(Y, row + (int64_t)col * ldy, row < m && col < k, re, im);
          p1[mb][nb][r] = re;
          p2[mb][nb][r] = T(0);
          p3[mb][nb][r] = re + im;
        }
    // Update from the blocks already solved: those left of J for R, right of
    // J for R^H.
    const int kbeg = CT ? cend : 0;
    const int kend = CT ? k : c0;
    if (kbeg < kend) {
      fetch_a(kbeg, kend);
      fetch_r(kbeg, c0);
    }
    for (int kc = kbeg; kc < kend; kc += TW_KC) {
      __syncthreads();
      stash();
      __syncthreads();
      if (kc + TW_KC < kend) {
        fetch_a(kc + TW_KC, kend);
        fetch_r(kc + TW_KC, c0);
      }
      if (active) mma(p1, p2, p3, Sub());
    }
    // The updated right-hand side goes back into the workgroup's rows of Y_J,
    // where the diagonal multiply reads it as its A operand.
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        acc_t re = p1[mb][nb], im = p2[mb][nb];
        if constexpr (CPLX) combine3m<PROD_ADD>(p1[mb][nb], p2[mb][nb], p3[mb][nb], re, im);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + nj0 + 16 * nb + li;
          if (row < m && col < k) store_elem<TS, CPLX>(Y, row + (int64_t)col * ldy, (TS)re[r], (TS)im[r]);
        }
      }
    // Workgroup-scope release/acquire: the stores above are visible to the
    // loads below.
    __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        p1[mb][nb] = Mfma<T>::zero();
        p2[mb][nb] = Mfma<T>::zero();
        p3[mb][nb] = Mfma<T>::zero();
      }
    const T *Dj = Dinv + (int64_t)J * TW_CB * TW_CB * NP;
    const int nbj = cend - c0;
    fetch_a(c0, cend);
    fetch_d(0, Dj);
    for (int kc = 0; kc < nbj; kc += TW_KC) {
      __syncthreads();
      stash();
      __syncthreads();
      if (kc + TW_KC < nbj) {
        fetch_a(c0 + kc + TW_KC, cend);
        fetch_d(kc + TW_KC, Dj);
      }
      // inv(R_JJ) is upper triangular: slice rows past the wave's last column
      // add nothing (lower triangular for R^H: rows before its first column).
      const bool contributes = CT ? (kc + TW_KC > nj0) : (kc < nj0 + 32);
      if (active && contributes) mma(p1, p2, p3, Add());
    }
    // Every wave has read its last slice of Y_J before any wave overwrites it.
    __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        acc_t re = p1[mb][nb], im = p2[mb][nb];
        if constexpr (CPLX) combine3m<PROD_ADD>(p1[mb][nb], p2[mb][nb], p3[mb][nb], re, im);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + nj0 + 16 * nb + li;
          if (row < m && col < k) store_elem<TS, CPLX>(Y, row + (int64_t)col * ldy, (TS)re[r], (TS)im[r]);
        }
      }
    // The solved block is read back by the next block's update.
    __syncthreads();
  }
}

template <typename TS, typename T, bool CPLX, bool CT = false>
static int launch_trsm(TS *Y, int64_t ldy, int64_t m, int k, const TS *R, int64_t ldr, T *Dinv,
                       hipStream_t stream) {
  const int nbk = (k + TW_CB - 1) / TW_CB;
  hipLaunchKernelGGL((trsm_block_inv_kernel<TS, T, CPLX>), dim3(nbk), dim3(TI_THREADS), 0, stream, R, ldr, k, Dinv);
  int st = check_launch("trsm_block_inv_kernel");
  if (st || m == 0) return st;
  const int64_t grid = cdiv64(m, TW_ROWS);
  if (grid > 0x7fffffff) return fail(FHIP_E_ARGUMENT, "fhip_trsm: too many rows (%lld)", (long long)m);
  hipLaunchKernelGGL((trsm_wide_kernel<TS, T, CPLX, CT>), dim3((unsigned)grid), dim3(TW_THREADS), 0, stream, Y,
                     ldy, m, k, R, ldr, (const T *)Dinv);
  return check_launch("trsm_wide_kernel");
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_trsm_workspace_bytes(size_t *nbytes, int64_t k, int compute) {
  if (!nbytes) return fail(FHIP_E_ARGUMENT, "fhip_trsm_workspace_bytes: null output pointer");
  if (!valid_type(compute)) return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runn: unknown type %d", compute);
  if (k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: negative dimension");
  *nbytes = (size_t)cdiv64(k, TW_CB) * TW_CB * TW_CB * eltype_size(compute);
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
