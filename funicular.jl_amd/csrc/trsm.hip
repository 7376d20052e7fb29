// Kernel (b): the right-upper triangular solve Y <- Y R^-1 that replaces
// rdiv_upper! (src/kernels.jl:108, semantics :90-99, called from rdiv_rows!
// src/panelblas.jl:157-169), and its conjugate-transpose twin Y <- Y R^-H
// for the QR adjoint.
//
// Design (gfx950). Rows are independent, so a workgroup owns TW_ROWS rows and
// walks the columns in blocks of TW_CB, left looking. With the diagonal
// inverses folded into the factor,
//     M_JJ = inv(R_JJ),   M_IJ = -R_IJ inv(R_JJ)  (I < J),
// block J of the solution is one product over the columns already at hand,
//     X_J = [Y_J, X_0, ..., X_{J-1}] * [M_JJ; M_0J; ...; M_{J-1,J}],
// so the whole solve is a chain of sums that never leaves the matrix cores'
// pipeline: the workgroup streams 16-column slices of its rows and of M
// through LDS, the loads of the slice after next in flight during the current
// slice's three-product (3M) complex steps, across block boundaries too. Each block
// reads its unsolved columns first and the solved ones after, so a slice in
// flight never races the block being written: the newest solved block is
// read several barriers after its stores. M is formed once per call in
// double: trsm_block_inv_kernel inverts the diagonal blocks (a right-looking
// sweep over the rows of the inverse, one barrier per row) and
// trsm_offdiag_kernel multiplies them into the blocks above. Per row, the
// workgroup reads sum_J TW_CB (J + 1) columns and writes k (4.5 k reads at
// k = 512 with 64-column blocks: 64 x (1 + 2 + ... + 8) = 2304 columns).
//
// Numerically this is blocked substitution with inverted diagonal blocks: the
// error of a block is governed by kappa of its 64 x 64 diagonal block, not by
// kappa(R).
#include "fhip_internal.hpp"

#include <mutex>
#include <type_traits>

namespace fhip {

#ifndef FHIP_TRSM_CB
#define FHIP_TRSM_CB 64
#endif
// Inner dimension per LDS slice for FP32 compute (a tuning knob; 32 measured
// slower than 16 at 629856 x 512).
#ifndef FHIP_TRSM_KC_FP32
#define FHIP_TRSM_KC_FP32 16
#endif
constexpr int TW_THREADS = 256;  // 4 waves, each a 32 x 32 tile of the block
constexpr int TW_CB = FHIP_TRSM_CB;
constexpr int TW_ROWS = 4096 / TW_CB;
// The solve Y R^-1 (trsm_wide_kernel) runs 8 waves, each a 32 x 16 tile of
// the block: 128 registers a lane, so two workgroups (four waves a SIMD)
// share a CU, against two waves a SIMD for 32 x 32 tiles at 234 registers.
// Its thread and row counts are tuning knobs: 1024 threads over 64 rows of
// 128-column blocks, or over 128 rows of 64-column blocks, hold one
// workgroup a CU and measured slower.
#ifndef FHIP_TRSM_WTHREADS
#define FHIP_TRSM_WTHREADS 512
#endif
#ifndef FHIP_TRSM_WROWS
#define FHIP_TRSM_WROWS 64
#endif
constexpr int TWW_THREADS = FHIP_TRSM_WTHREADS;
constexpr int TWW_ROWS = FHIP_TRSM_WROWS;  // rows per workgroup of trsm_wide_kernel
constexpr int TWW_WAVES_EU = TWW_THREADS == 256 ? 2 : 4;
#ifndef FHIP_TRSM_SWIZZLE
#define FHIP_TRSM_SWIZZLE 1
#endif
constexpr int TI_THREADS = 8 * TW_CB;  // block inverse: 8 row groups x TW_CB columns
constexpr int TI_RPT = TW_CB / 8;      // rows of the inverse per thread
constexpr int TI_CHUNK = 16;           // columns of R staged in LDS at a time
static_assert(TW_CB == 32 || TW_CB == 64 || TW_CB == 128, "column block of 32, 64 or 128");

#ifndef FHIP_TRSM_KC_FP64
#define FHIP_TRSM_KC_FP64 16
#endif
template <typename T>
struct TrsmSlice {
  static constexpr int KC = sizeof(T) == 4 ? FHIP_TRSM_KC_FP32 : FHIP_TRSM_KC_FP64;
  static constexpr int APT = KC * TW_ROWS / TW_THREADS;  // A elements per thread per slice
  static constexpr int BPT = KC * TW_CB / TW_THREADS;    // B elements per thread per slice
};

// M is kp x kp, kp = k rounded up to TW_CB, column major.
inline int64_t trsm_kp(int64_t k) { return cdiv64(k, TW_CB) * TW_CB; }

// Inverse of each TW_CB x TW_CB diagonal block of R (the last one cut to
// what is left of k, zero padded), written into the diagonal blocks of M,
// conjugate transposed when CT. Computed in double. W = inv(R_JJ) solves
// R_JJ W = I from the bottom row up: once row l of W is known, its multiple
// R[i, l] W[l, :] is taken off every row i < l. Thread (g, j) keeps column j
// of rows g, g + 8, ... in registers; R comes through LDS sixteen columns at
// a time.
template <typename TS, typename T, bool CPLX, bool CT>
__global__ __launch_bounds__(TI_THREADS) void trsm_block_inv_kernel(const TS *__restrict__ R, int64_t ldr, int k,
                                                                     T *__restrict__ M, int64_t kp) {
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
#pragma unroll
  for (int q = 0; q < TI_RPT; ++q) {
    const int i = g + 8 * q;
    const bool keep = i <= j && j < nb;
    // Element (i, j) of inv(R_JJ), or element (j, i) of its conjugate transpose.
    const int64_t at = CT ? (int64_t)(c0 + j) + (int64_t)(c0 + i) * kp : (int64_t)(c0 + i) + (int64_t)(c0 + j) * kp;
    M[at * NP] = keep ? (T)br[q] : T(0);
    if (CPLX) M[at * NP + 1] = keep ? (T)(CT ? -bi[q] : bi[q]) : T(0);
  }
}

// The off-diagonal blocks of M: M_IJ = -F_IJ M_JJ with F = R (I < J) or
// F = R^H (I > J, when CT), one workgroup per block, in double, the factors
// staged in LDS 32 inner columns at a time.
template <typename TS, typename T, bool CPLX, bool CT>
__global__ __launch_bounds__(256) void trsm_offdiag_kernel(const TS *__restrict__ R, int64_t ldr, int k,
                                                           T *__restrict__ M, int64_t kp) {
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int B = TW_CB;
  constexpr int LC = 32 < B ? 32 : B;             // inner columns per pass
  constexpr int ROWS = B / (256 / B);             // output rows per thread
  __shared__ double f_re[B][LC + 1], f_im[B][LC + 1];  // F_IJ[i][l0 + l]
  __shared__ double d_re[LC][B + 1], d_im[LC][B + 1];  // M_JJ[l0 + l][j]
  // Pair p -> (I, J): I < J, J = 1, 2, ..., pairs of one J consecutive.
  int p = blockIdx.x, J = 1;
  while (p >= J) {
    p -= J;
    ++J;
  }
  const int nbk = (k + B - 1) / B;
  int I = p;
  if (CT) {
    // Mirror: I > J.
    const int Jm = nbk - 1 - J, Im = nbk - 1 - I;
    I = Im;
    J = Jm;
  }
  const int cI = I * B, cJ = J * B;
  const int t = threadIdx.x;
  const int j = t % B;
  const int g = t / B;
  double sr[ROWS], si[ROWS];
  for (int q = 0; q < ROWS; ++q) sr[q] = si[q] = 0.0;
  for (int l0 = 0; l0 < B; l0 += LC) {
    __syncthreads();
    for (int e = t; e < B * LC; e += 256) {
      const int i = e % B, l = e / B;
      double fr = 0.0, fi = 0.0;
      // F_IJ[i][l0 + l]: R[cI + i, cJ + l0 + l], or conj(R[cJ + l0 + l, cI + i]) for R^H.
      const int ri = CT ? cJ + l0 + l : cI + i;
      const int rl = CT ? cI + i : cJ + l0 + l;
      if (ri < k && rl < k) {
        const int64_t idx = (int64_t)ri + (int64_t)rl * ldr;
        fr = (double)R[CPLX ? 2 * idx : idx];
        fi = CPLX ? (double)R[2 * idx + 1] : 0.0;
        if (CT) fi = -fi;
      }
      f_re[i][l] = fr;
      f_im[i][l] = fi;
      // M_JJ[l0 + l][i] (here i runs over the columns of M_JJ).
      const int64_t at = (int64_t)(cJ + l0 + l) + (int64_t)(cJ + i) * kp;
      d_re[l][i] = (double)M[at * NP];
      d_im[l][i] = CPLX ? (double)M[at * NP + 1] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < ROWS; ++q) {
      const int i = g + (256 / B) * q;
      for (int l = 0; l < LC; ++l) {
        const double ar = f_re[i][l], ai = f_im[i][l], br = d_re[l][j], bi = d_im[l][j];
        sr[q] += ar * br - ai * bi;
        si[q] += ar * bi + ai * br;
      }
    }
  }
  for (int q = 0; q < ROWS; ++q) {
    const int i = g + (256 / B) * q;
    const int64_t at = (int64_t)(cI + i) + (int64_t)(cJ + j) * kp;
    M[at * NP] = (T)(-sr[q]);
    if (CPLX) M[at * NP + 1] = (T)(-si[q]);
  }
}

// The solve. CT = false solves X R = Y (blocks left to right); CT = true
// solves X R^H = Y, whose factor is lower triangular (blocks right to left).
// Each wave computes a 32 x 32 piece of the workgroup's TW_ROWS x TW_CB block
// as 2 x 2 matrix-core tiles.
//
// Thread t stages rows t % TW_ROWS of columns t / TW_ROWS + YCS e of each Y
// slice and rows t % KC of columns t / KC + MCS e of each M slice, so the
// address of every element it loads is its base plus a wave-uniform offset
// (kc and the element index times the column stride): one 64-bit add per
// load, no 64-bit multiply and no per-element branch. The slice schedule is
// kept in scalars updated once per slice.
//
// trsm_wide_short_kernel issues each slice's loads one slice ahead; it fits
// the registers of two waves per SIMD for both directions, and serves R^H.
// trsm_wide_kernel (below) issues them two slices ahead.
template <typename TS, typename T, bool CPLX, bool CT>
__global__ __launch_bounds__(TW_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void trsm_wide_short_kernel(
    TS *__restrict__ Y, int64_t ldy, int64_t m, int k, const T *__restrict__ M, int64_t kp) {
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int TW_KC = TrsmSlice<T>::KC;
  constexpr int TW_APT = TrsmSlice<T>::APT;
  constexpr int TW_BPT = TrsmSlice<T>::BPT;
  constexpr int SWZ = !(CPLX && FHIP_TRSM_SWIZZLE) ? 0
                      : sizeof(T) == 8                                     ? 1
                      : (TW_ROWS == 64 && TW_CB == 64 && TW_KC == 16)      ? 2
                                                                           : 0;
  constexpr int TW_LDA = SWZ ? TW_ROWS : TW_ROWS + 4;
  constexpr int TW_LDB = SWZ ? TW_CB : TW_CB + 4;
  constexpr int YCS = TW_THREADS / TW_ROWS;  // column step between a thread's Y elements
  constexpr int MCS = TW_THREADS / TW_KC;    // column step between a thread's M elements
  static_assert(TW_THREADS % TW_ROWS == 0 && TW_THREADS % TW_KC == 0, "staging layout");
  static_assert(YCS * TW_APT == TW_KC && MCS * TW_BPT == TW_CB, "staging covers the slice");
  __shared__ T As[NP][TW_KC * TW_LDA];
  __shared__ T Bs[NP][TW_KC * TW_LDB];
  typedef typename Mfma<T>::acc acc_t;
  typedef typename Pair<TS>::type PS;
  typedef typename Pair<T>::type PT;

  const int64_t row0 = (int64_t)blockIdx.x * TW_ROWS;
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  constexpr int WAVE_ROWS = TW_ROWS / 32;
  const int mi0 = 32 * (w % WAVE_ROWS);
  const int nj0 = 32 * (w / WAVE_ROWS);
  const int nbk = (k + TW_CB - 1) / TW_CB;

  // Per-thread bases. A row past m reads row 0 (its results are never
  // stored); a column past k is zeroed when staged.
  const int yr = t % TW_ROWS, yk = t / TW_ROWS;
  const int mk = t % TW_KC, mc = t / TW_KC;
  const int64_t row = row0 + yr;
  const bool row_ok = row < m;
  const TS *ybase = Y + (int64_t)NP * ((row_ok ? row : 0) + (int64_t)yk * ldy);
  const T *mbase = M + (int64_t)NP * (mk + (int64_t)mc * kp);
  const int64_t ystride = (int64_t)NP * ldy;  // one column of Y
  const int64_t mstride = (int64_t)NP * kp;   // one column of M

  PS ya[TW_APT];
  PT ma[TW_BPT];
  bool oka[TW_APT];
  // Slice at column kc of block J (first column c0): columns [kc, kc + KC)
  // of the workgroup's rows of Y and rows [kc, kc + KC) x the block's
  // columns of M. Every M load stays inside the kp x kp factor, whose padding
  // past k is zero.
  auto fetch = [&](int c0, int kc) {
    const TS *yp = ybase + (int64_t)kc * ystride;
#pragma unroll
    for (int e = 0; e < TW_APT; ++e) {
      const bool ok = row_ok && kc + yk + YCS * e < k;
      oka[e] = ok;
      const TS *p = ok ? yp + (int64_t)(YCS * e) * ystride : Y;
      if constexpr (CPLX) {
        ya[e] = *reinterpret_cast<const PS *>(p);
      } else {
        ya[e].x = *p;
        ya[e].y = TS(0);
      }
    }
    const T *mp = mbase + (int64_t)NP * kc + (int64_t)c0 * mstride;
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const T *p = mp + (int64_t)(MCS * e) * mstride;
      if constexpr (CPLX) {
        ma[e] = *reinterpret_cast<const PT *>(p);
      } else {
        ma[e].x = *p;
        ma[e].y = T(0);
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int e = 0; e < TW_APT; ++e) {
      const int kk = yk + YCS * e;
      const int at = kk * TW_LDA + (yr ^ swizzle_of<SWZ>(kk));
      T re, im;
      from_raw(ya[e], oka[e], re, im);
      As[0][at] = re;
      if (CPLX) As[NP - 1][at] = im;
    }
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const int c = mc + MCS * e;
      const int at = mk * TW_LDB + (c ^ swizzle_of<SWZ>(mk));
      Bs[0][at] = ma[e].x;
      if (CPLX) Bs[NP - 1][at] = ma[e].y;
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

  // The slice schedule: block J = the step's block, its own slices first
  // (from c0), then the solved ones (from 0 for R, from c0 + CB for R^H).
  auto own_of = [&](int c0) { return ((k - c0 < TW_CB ? k - c0 : TW_CB) + TW_KC - 1) / TW_KC; };
  auto solved_start = [&](int c0) { return CT ? c0 + TW_CB : 0; };
  auto count_of = [&](int c0) {
    const int solved = CT ? k - (c0 + TW_CB < k ? c0 + TW_CB : k) : c0;
    return own_of(c0) + (solved + TW_KC - 1) / TW_KC;
  };
  int step = 0;
  int c0 = (CT ? nbk - 1 : 0) * TW_CB;
  int own = own_of(c0), n = count_of(c0);
  int idx = 0;
  int kc = c0;
  fetch(c0, kc);
  while (true) {
    __syncthreads();
    stash();
    __syncthreads();
    // Prefetch the next slice, crossing into the next block.
    if (idx + 1 < n) {
      const int kcn = idx + 1 < own ? kc + TW_KC : (idx + 1 == own ? solved_start(c0) : kc + TW_KC);
      fetch(c0, kcn);
    } else if (step + 1 < nbk) {
      const int c0n = (CT ? nbk - 2 - step : step + 1) * TW_CB;
      fetch(c0n, c0n);
    }
    bool contributes = c0 + nj0 < k;
    if (idx < own) {
      const int kl = kc - c0;
      contributes = contributes && (CT ? (kl + TW_KC > nj0) : (kl < nj0 + 32));
    }
    if (contributes) {
      if constexpr (CPLX)
        mma_steps_3m<T, 2, 2, PROD_ADD, SWZ>(As[0], As[NP - 1], TW_LDA, mi0, Bs[0], Bs[NP - 1], TW_LDB, nj0, 0,
                                             TW_KC / 4, 1, p1, p2, p3);
      else
        mma_steps<T, false, 2, 2, PROD_ADD>(As[0], As[0], TW_LDA, mi0, Bs[0], Bs[0], TW_LDB, nj0, 0, TW_KC / 4,
                                            1, p1, p2);
    }
    if (++idx < n) {
      kc = idx < own ? kc + TW_KC : (idx == own ? solved_start(c0) : kc + TW_KC);
      continue;
    }
    // Block J is solved. Every wave has staged its last slice of Y_J (the
    // barriers above), so the stores cannot race a load of the unsolved
    // columns; the next blocks read X_J only after further barriers.
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        acc_t re = p1[mb][nb], im = p2[mb][nb];
        if constexpr (CPLX) combine3m<PROD_ADD>(p1[mb][nb], p2[mb][nb], p3[mb][nb], re, im);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t orow = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + nj0 + 16 * nb + li;
          if (orow < m && col < k) store_elem<TS, CPLX>(Y, orow + (int64_t)col * ldy, (TS)re[r], (TS)im[r]);
        }
      }
    if (++step == nbk) break;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        p1[mb][nb] = Mfma<T>::zero();
        p2[mb][nb] = Mfma<T>::zero();
        p3[mb][nb] = Mfma<T>::zero();
      }
    c0 = (CT ? nbk - 1 - step : step) * TW_CB;
    own = own_of(c0);
    n = count_of(c0);
    idx = 0;
    kc = c0;
  }
}

// The solve with each slice's global loads issued two slices ahead: two sets
// of staging registers alternate, so a slice's loads have two slices of
// products to land in instead of one (the reads of solved columns come back
// from HBM). The loop body is unrolled twice so that each set is a fixed set
// of registers. Within a diagonal block, a 16-row slice of the triangular
// M_JJ meets each 16-column tile of a wave as a zero block, a triangle or a
// full block; the products with zero blocks are left out.
//
// With BUF (complex storage) the slices are read with raw buffer loads. A
// wave stages one column of the block per element (its 64 lanes are the
// block's 64 rows), so each of its Y loads gets a descriptor of its own,
// built with scalar arithmetic: based at the column's row row0 and ranging
// over the block's rows of that column that lie below m, or over nothing for a
// column past k. The lane's offset, its row times the element size, is fixed
// for the whole solve; rows past m and columns past k read as zero with no
// per-lane select or 64-bit address arithmetic, and nothing outside the
// block's own rows is read, whatever its leading dimension. Every vector instruction
// beside the matrix-core products costs matrix-core time on gfx950 (about
// five cycles each), so the selects and 64-bit address updates this drops are
// worth taking out of each slice.
template <typename TS, typename T, bool CPLX, bool CT, bool BUF = false>
__global__ __launch_bounds__(TWW_THREADS) __attribute__((amdgpu_waves_per_eu(TWW_WAVES_EU))) void trsm_wide_kernel(
    TS *__restrict__ Y, int64_t ldy, int64_t m, int k, const T *__restrict__ M, int64_t kp) {
  static_assert(!BUF || CPLX, "buffer loads are for complex storage");
  static_assert(!BUF || TWW_ROWS == 64, "buffer loads take one column of the block per wave");
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int TW_KC = TrsmSlice<T>::KC;
  constexpr int TW_APT = TW_KC * TWW_ROWS / TWW_THREADS;
  constexpr int TW_BPT = TW_KC * TW_CB / TWW_THREADS;
  constexpr int WNB = TWW_THREADS == 256 ? 2 : 1;  // 16-column tiles per wave
  constexpr int SWZ = !(CPLX && FHIP_TRSM_SWIZZLE) ? 0
                      : sizeof(T) == 8                                     ? 1
                      : (TWW_ROWS == 64 && TW_CB == 64 && TW_KC == 16)      ? 2
                                                                           : 0;
  constexpr int TW_LDA = SWZ ? TWW_ROWS : TWW_ROWS + 4;
  constexpr int TW_LDB = SWZ ? TW_CB : TW_CB + 4;
  constexpr int YCS = TWW_THREADS / TWW_ROWS;
  constexpr int MCS = TWW_THREADS / TW_KC;
  static_assert(TWW_THREADS % TWW_ROWS == 0 && TWW_THREADS % TW_KC == 0, "staging layout");
  static_assert((TWW_THREADS / 64) / (TWW_ROWS / 32) * 16 * WNB == TW_CB, "the waves tile the block");
  static_assert(YCS * TW_APT == TW_KC && MCS * TW_BPT == TW_CB, "staging covers the slice");
  __shared__ T As[NP][TW_KC * TW_LDA];
  __shared__ T Bs[NP][TW_KC * TW_LDB];
  typedef typename Mfma<T>::acc acc_t;
  typedef typename Pair<TS>::type PS;
  typedef typename Pair<T>::type PT;

  const int64_t row0 = (int64_t)blockIdx.x * TWW_ROWS;
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  constexpr int WAVE_ROWS = TWW_ROWS / 32;
  // The wave's tile offsets as scalars: the per-slice test of whether its
  // column tile takes part then runs on the scalar unit.
  const int mi0 = __builtin_amdgcn_readfirstlane(32 * (w % WAVE_ROWS));
  const int nj0 = __builtin_amdgcn_readfirstlane(16 * WNB * (w / WAVE_ROWS));
  const int nbk = (k + TW_CB - 1) / TW_CB;

  const int yr = t % TWW_ROWS, yk = t / TWW_ROWS;
  const int mk = t % TW_KC, mc = t / TW_KC;
  const int64_t row = row0 + yr;
  const bool row_ok = row < m;
  const TS *ybase = Y + (int64_t)NP * ((row_ok ? row : 0) + (int64_t)yk * ldy);
  const T *mbase = M + (int64_t)NP * (mk + (int64_t)mc * kp);
  const int64_t ystride = (int64_t)NP * ldy;
  const int64_t mstride = (int64_t)NP * kp;

  PS ya[2][TW_APT];
  PT ma[2][TW_BPT];
  bool oka[2][TW_APT];
  // Buffer loads (BUF): the wave's column within a slice (wave-uniform),
  // the lane's byte offset in a column, the bytes of a column's rows below m,
  // and the offsets of the thread's factor elements.
  const int wcol = __builtin_amdgcn_readfirstlane(yk);
  const uint32_t yoff = (uint32_t)(yr * sizeof(PS));
  const uint32_t col_bytes = (uint32_t)((m - row0 < TWW_ROWS ? m - row0 : TWW_ROWS) * (int64_t)sizeof(PS));
  uint32_t moff[TW_BPT];
#pragma unroll
  for (int e = 0; e < TW_BPT; ++e) moff[e] = (uint32_t)((mk + (int64_t)(mc + MCS * e) * kp) * (int64_t)sizeof(PT));
  auto fetch = [&](auto set, int c0, int kc) {
    constexpr int S = decltype(set)::value;
    if constexpr (BUF) {
#pragma unroll
      for (int e = 0; e < TW_APT; ++e) {
        const int col = kc + wcol + YCS * e;
        const auto ry = buffer_rsrc(Y + (int64_t)NP * (row0 + (int64_t)col * ldy), col < k ? col_bytes : 0u);
        ya[S][e] = buffer_load_pair<PS>(ry, yoff);
      }
      const int64_t mfirst = (int64_t)kc + (int64_t)c0 * kp;
      const auto rm = buffer_rsrc(M + (int64_t)NP * mfirst, (uint32_t)((kp * kp - mfirst) * (int64_t)sizeof(PT)));
#pragma unroll
      for (int e = 0; e < TW_BPT; ++e) ma[S][e] = buffer_load_pair<PT>(rm, moff[e]);
      return;
    }
    const TS *yp = ybase + (int64_t)kc * ystride;
#pragma unroll
    for (int e = 0; e < TW_APT; ++e) {
      const bool ok = row_ok && kc + yk + YCS * e < k;
      oka[S][e] = ok;
      const TS *p = ok ? yp + (int64_t)(YCS * e) * ystride : Y;
      if constexpr (CPLX) {
        ya[S][e] = *reinterpret_cast<const PS *>(p);
      } else {
        ya[S][e].x = *p;
        ya[S][e].y = TS(0);
      }
    }
    const T *mp = mbase + (int64_t)NP * kc + (int64_t)c0 * mstride;
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const T *p = mp + (int64_t)(MCS * e) * mstride;
      if constexpr (CPLX) {
        ma[S][e] = *reinterpret_cast<const PT *>(p);
      } else {
        ma[S][e].x = *p;
        ma[S][e].y = T(0);
      }
    }
  };
  auto stash = [&](auto set) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int e = 0; e < TW_APT; ++e) {
      const int kk = yk + YCS * e;
      const int at = kk * TW_LDA + (yr ^ swizzle_of<SWZ>(kk));
      T re, im;
      if constexpr (BUF) {
        re = (T)ya[S][e].x;
        im = (T)ya[S][e].y;
      } else {
        from_raw(ya[S][e], oka[S][e], re, im);
      }
      As[0][at] = re;
      if (CPLX) As[NP - 1][at] = im;
    }
#pragma unroll
    for (int e = 0; e < TW_BPT; ++e) {
      const int c = mc + MCS * e;
      const int at = mk * TW_LDB + (c ^ swizzle_of<SWZ>(mk));
      Bs[0][at] = ma[S][e].x;
      if (CPLX) Bs[NP - 1][at] = ma[S][e].y;
    }
  };

  acc_t p1[2][WNB], p2[2][WNB], p3[2][WNB];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < WNB; ++nb) {
      p1[mb][nb] = Mfma<T>::zero();
      p2[mb][nb] = Mfma<T>::zero();
      p3[mb][nb] = Mfma<T>::zero();
    }

  // Slice iterators: the slice being multiplied and the next one to fetch.
  // A block's own slices come first. When a block has solved slices after a
  // single slice of its own, a second, empty own slice is staged (columns
  // past k load as zeros; no products are formed from it): the first solved
  // slice is then fetched two slices after the previous block was stored,
  // never before.
  struct It {
    int step, c0, real, own, n, idx, kc;
    bool valid;
  };
  auto start = [&](It &it, int step) {
    it.step = step;
    it.c0 = (CT ? nbk - 1 - step : step) * TW_CB;
    it.real = ((k - it.c0 < TW_CB ? k - it.c0 : TW_CB) + TW_KC - 1) / TW_KC;
    const int solved = CT ? k - (it.c0 + TW_CB < k ? it.c0 + TW_CB : k) : it.c0;
    const int nsolved = (solved + TW_KC - 1) / TW_KC;
    it.own = (it.real < 2 && nsolved > 0) ? 2 : it.real;
    it.n = it.own + nsolved;
    it.idx = 0;
    it.kc = it.c0;
    it.valid = step < nbk;
  };
  auto next = [&](It &it) {
    if (++it.idx < it.n) {
      it.kc = it.idx < it.own ? it.kc + TW_KC : (it.idx == it.own ? (CT ? it.c0 + TW_CB : 0) : it.kc + TW_KC);
    } else {
      start(it, it.step + 1);
    }
  };
  It cur, pf;
  start(cur, 0);
  pf = cur;
  fetch(std::integral_constant<int, 0>(), pf.c0, pf.kc);
  next(pf);
  if (pf.valid) {
    fetch(std::integral_constant<int, 1>(), pf.c0, pf.kc);
    next(pf);
  }
  // One slice: stage set S, refill it two slices ahead, multiply, and store
  // a finished block. Returns false after the last slice.
  auto slice = [&](auto set) {
    __syncthreads();
    stash(set);
    __syncthreads();
    if (pf.valid) {
      fetch(set, pf.c0, pf.kc);
      next(pf);
    }
    const bool last = cur.idx + 1 >= cur.n;
    const int c0 = cur.c0;
    bool contributes = c0 + nj0 < k && !(cur.idx >= cur.real && cur.idx < cur.own);
    constexpr unsigned FULL = (1u << WNB) - 1u;
    unsigned nb_mask = FULL;
    if (cur.idx < cur.own) {
      const int kl = cur.kc - c0;
      contributes = contributes && (CT ? (kl + TW_KC > nj0) : (kl < nj0 + 16 * WNB));
      if constexpr (TW_KC == 16) {
        const int rb = kl >> 4, cb0 = nj0 >> 4;
        if constexpr (WNB == 2)
          nb_mask = CT ? ((rb >= cb0 ? 1u : 0u) | (rb >= cb0 + 1 ? 2u : 0u))
                       : ((rb <= cb0 ? 1u : 0u) | (rb <= cb0 + 1 ? 2u : 0u));
        else
          nb_mask = CT ? (rb >= cb0 ? 1u : 0u) : (rb <= cb0 ? 1u : 0u);
      }
    }
    // With one 16-column tile a wave, a tile's product is either whole or
    // left out: one product path.
    if constexpr (WNB == 1) contributes = contributes && nb_mask == FULL;
    if (contributes && nb_mask == FULL) {
      if constexpr (CPLX)
        mma_steps_3m<T, 2, WNB, PROD_ADD, SWZ>(As[0], As[NP - 1], TW_LDA, mi0, Bs[0], Bs[NP - 1], TW_LDB, nj0, 0,
                                             TW_KC / 4, 1, p1, p2, p3);
      else
        mma_steps<T, false, 2, WNB, PROD_ADD>(As[0], As[0], TW_LDA, mi0, Bs[0], Bs[0], TW_LDB, nj0, 0, TW_KC / 4,
                                            1, p1, p2);
    } else if (WNB > 1 && contributes) {
      if constexpr (CPLX)
        mma_steps_3m<T, 2, WNB, PROD_ADD, SWZ>(As[0], As[NP - 1], TW_LDA, mi0, Bs[0], Bs[NP - 1], TW_LDB, nj0, 0,
                                             TW_KC / 4, 1, p1, p2, p3, nb_mask);
      else
        mma_steps<T, false, 2, WNB, PROD_ADD>(As[0], As[0], TW_LDA, mi0, Bs[0], Bs[0], TW_LDB, nj0, 0, TW_KC / 4,
                                            1, p1, p2);
    }
    if (!last) {
      next(cur);
      return true;
    }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < WNB; ++nb) {
        acc_t re = p1[mb][nb], im = p2[mb][nb];
        if constexpr (CPLX) combine3m<PROD_ADD>(p1[mb][nb], p2[mb][nb], p3[mb][nb], re, im);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t orow = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + nj0 + 16 * nb + li;
          if (orow < m && col < k) store_elem<TS, CPLX>(Y, orow + (int64_t)col * ldy, (TS)re[r], (TS)im[r]);
        }
        p1[mb][nb] = Mfma<T>::zero();
        p2[mb][nb] = Mfma<T>::zero();
        p3[mb][nb] = Mfma<T>::zero();
      }
    next(cur);
    return cur.valid;
  };
  while (slice(std::integral_constant<int, 0>()) && slice(std::integral_constant<int, 1>())) {
  }
}

// The solve proper, with the folded factor M already in the workspace.
template <typename TS, typename T, bool CPLX, bool CT = false>
static int launch_wide(TS *Y, int64_t ldy, int64_t m, int k, const T *M, hipStream_t stream) {
  if (m == 0) return FHIP_OK;
  const int64_t kp = trsm_kp(k);
  const int64_t grid = cdiv64(m, CT ? TW_ROWS : TWW_ROWS);
  if (grid > 0x7fffffff) return fail(FHIP_E_ARGUMENT, "fhip_trsm: too many rows (%lld)", (long long)m);
  // The buffer path needs the folded factor inside a 32-bit byte range (Y's
  // descriptors each span one column of the block's rows).
#ifndef FHIP_TRSM_BUF
#define FHIP_TRSM_BUF 1
#endif
  const bool buf = FHIP_TRSM_BUF && CPLX && TWW_ROWS == 64 && kp * kp * 2 * (int64_t)sizeof(T) < (int64_t(1) << 32);
  if constexpr (CT) {
    hipLaunchKernelGGL((trsm_wide_short_kernel<TS, T, CPLX, CT>), dim3((unsigned)grid), dim3(TW_THREADS), 0, stream,
                       Y, ldy, m, k, M, kp);
  } else if constexpr (CPLX) {
    if (buf)
      hipLaunchKernelGGL((trsm_wide_kernel<TS, T, CPLX, CT, true>), dim3((unsigned)grid), dim3(TWW_THREADS), 0,
                         stream, Y, ldy, m, k, M, kp);
    else
      hipLaunchKernelGGL((trsm_wide_kernel<TS, T, CPLX, CT>), dim3((unsigned)grid), dim3(TWW_THREADS), 0, stream,
                         Y, ldy, m, k, M, kp);
  } else {
    hipLaunchKernelGGL((trsm_wide_kernel<TS, T, CPLX, CT>), dim3((unsigned)grid), dim3(TWW_THREADS), 0, stream, Y,
                       ldy, m, k, M, kp);
  }
  return check_launch("trsm_wide_kernel");
}

template <typename TS, typename T, bool CPLX, bool CT = false>
static int launch_trsm(TS *Y, int64_t ldy, int64_t m, int k, const TS *R, int64_t ldr, T *M, hipStream_t stream) {
  const int nbk = (k + TW_CB - 1) / TW_CB;
  const int64_t kp = trsm_kp(k);
  hipLaunchKernelGGL((trsm_block_inv_kernel<TS, T, CPLX, CT>), dim3(nbk), dim3(TI_THREADS), 0, stream, R, ldr, k, M,
                     kp);
  int st = check_launch("trsm_block_inv_kernel");
  if (st) return st;
  const int npairs = nbk * (nbk - 1) / 2;
  if (npairs > 0) {
    hipLaunchKernelGGL((trsm_offdiag_kernel<TS, T, CPLX, CT>), dim3(npairs), dim3(256), 0, stream, R, ldr, k, M, kp);
    st = check_launch("trsm_offdiag_kernel");
    if (st) return st;
  }
  return launch_wide<TS, T, CPLX, CT>(Y, ldy, m, k, M, stream);
}

// Checks shared by fhip_trsm_runn and fhip_trsm_runn_noalloc, then the solve
// with the folded factor in M (m_bytes long). `name` labels the messages.
static int trsm_runn_checked(const char *name, void *Y, int64_t ldy, int64_t m, int64_t k, const void *R,
                             int64_t ldr, int type, int compute, void *M, size_t m_bytes, hipStream_t s) {
  if (!valid_type(type) || !valid_type(compute))
    return fail(FHIP_E_UNSUPPORTED, "%s: unknown type %d / %d", name, type, compute);
  if (!(compute == type || (type == FHIP_C64 && compute == FHIP_C32)))
    return fail(FHIP_E_UNSUPPORTED, "%s: no kernel for storage type %d with compute type %d", name, type, compute);
  if (m < 0 || k < 0) return fail(FHIP_E_ARGUMENT, "%s: negative dimension", name);
  if (k == 0) return FHIP_OK;
  if (k > (1 << 16)) return fail(FHIP_E_ARGUMENT, "%s: too many columns", name);
  if (!Y || !R) return fail(FHIP_E_ARGUMENT, "%s: null matrix pointer", name);
  if (ldr < k || (m > 0 && ldy < m))
    return fail(FHIP_E_ARGUMENT, "%s: leading dimension smaller than the matrix", name);
  size_t need = 0;
  fhip_trsm_workspace_bytes(&need, k, compute);
  if (!M || m_bytes < need)
    return fail(FHIP_E_WORKSPACE, "%s: workspace of %zu bytes is smaller than the %zu needed", name, m_bytes, need);
  const int ik = (int)k;
  if (type == FHIP_C64 && compute == FHIP_C32)
    return launch_trsm<double, float, true>((double *)Y, ldy, m, ik, (const double *)R, ldr, (float *)M, s);
  switch (type) {
    case FHIP_C64:
      return launch_trsm<double, double, true>((double *)Y, ldy, m, ik, (const double *)R, ldr, (double *)M, s);
    case FHIP_C32:
      return launch_trsm<float, float, true>((float *)Y, ldy, m, ik, (const float *)R, ldr, (float *)M, s);
    case FHIP_F64:
      return launch_trsm<double, double, false>((double *)Y, ldy, m, ik, (const double *)R, ldr, (double *)M, s);
    case FHIP_F32:
      return launch_trsm<float, float, false>((float *)Y, ldy, m, ik, (const float *)R, ldr, (float *)M, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "%s: unknown type %d", name, type);
}

// The folded factor of fhip_trsm_runn_noalloc lives here: a device array of
// fixed size, one copy per GPU, held until the process ends. 16 MiB is the
// factor of k = 1024 in FP64 complex (kp^2 x 16 bytes). Host code must never
// pass the array's name to a kernel or a copy: it is a host-side stand-in,
// not a device address. scratch_address() resolves the address on the
// current device once and keeps it.
constexpr size_t TRSM_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[TRSM_SCRATCH_BYTES];

constexpr int TRSM_SCRATCH_DEVICES = 64;
static std::mutex g_scratch_lock;
static void *g_scratch_ptr[TRSM_SCRATCH_DEVICES];

static int scratch_address(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= TRSM_SCRATCH_DEVICES)
    return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn_noalloc: device index %d out of range", dev);
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(trsm scratch)");
    if (!p) return fail(FHIP_E_HIP, "hipGetSymbolAddress(trsm scratch) returned a null address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FHIP_OK;
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_trsm_workspace_bytes(size_t *nbytes, int64_t k, int compute) {
  if (!nbytes) return fail(FHIP_E_ARGUMENT, "fhip_trsm_workspace_bytes: null output pointer");
  if (!valid_type(compute)) return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runn: unknown type %d", compute);
  if (k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: negative dimension");
  if (k > (1 << 16)) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: too many columns (%lld)", (long long)k);
  const int64_t kp = trsm_kp(k);
  *nbytes = (size_t)kp * (size_t)kp * eltype_size(compute);
  return FHIP_OK;
}

extern "C" int fhip_trsm_runn(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr,
                              int type, int compute, void *work, size_t work_bytes, void *stream) {
  return trsm_runn_checked("fhip_trsm_runn", Y, ldy, m, k, R, ldr, type, compute, work, work_bytes,
                           (hipStream_t)stream);
}

extern "C" int fhip_trsm_runn_noalloc(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr,
                                      int type, int compute, void *stream) {
  if (valid_type(compute) && k >= 0 && k <= (1 << 16)) {
    size_t need = 0;
    fhip_trsm_workspace_bytes(&need, k, compute);
    if (need > TRSM_SCRATCH_BYTES)
      return fail(FHIP_E_ARGUMENT,
                  "fhip_trsm_runn_noalloc: k = %lld needs a %zu-byte folded factor, past the %zu bytes of the "
                  "library's fixed device scratch (k <= 1024 for complex FP64); use fhip_trsm_runn with a workspace",
                  (long long)k, need, TRSM_SCRATCH_BYTES);
  }
  void *M = nullptr;
  // Resolve the scratch only for a call that will launch: argument errors
  // are reported without touching a device.
  const bool launches = valid_type(type) && valid_type(compute) && k > 0 && k <= (1 << 16) && m >= 0 && Y && R &&
                        ldr >= k && (m == 0 || ldy >= m);
  if (launches) {
    int st = scratch_address(&M);
    if (st) return st;
  }
  return trsm_runn_checked("fhip_trsm_runn_noalloc", Y, ldy, m, k, R, ldr, type, compute, M,
                           launches ? TRSM_SCRATCH_BYTES : 0, (hipStream_t)stream);
}

extern "C" int fhip_trsm_runc(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr, int type,
                              void *work, size_t work_bytes, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runc: unknown type %d", type);
  if (m < 0 || k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runc: negative dimension");
  if (k == 0) return FHIP_OK;
  if (k > (1 << 16)) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runc: too many columns");
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

extern "C" int fhip_trsm_apply(void *Y, int64_t ldy, int64_t m, int64_t k, int type, int compute, int adjoint,
                               const void *work, size_t work_bytes, void *stream) {
  if (!valid_type(type) || !valid_type(compute))
    return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_apply: unknown type %d / %d", type, compute);
  if (!(compute == type || (type == FHIP_C64 && compute == FHIP_C32 && !adjoint)))
    return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_apply: no kernel for storage type %d with compute type %d", type,
                compute);
  if (m < 0 || k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_apply: negative dimension");
  if (k == 0 || m == 0) return FHIP_OK;
  if (k > (1 << 16)) return fail(FHIP_E_ARGUMENT, "fhip_trsm_apply: too many columns");
  if (!Y) return fail(FHIP_E_ARGUMENT, "fhip_trsm_apply: null matrix pointer");
  if (ldy < m) return fail(FHIP_E_ARGUMENT, "fhip_trsm_apply: leading dimension smaller than the matrix");
  size_t need = 0;
  fhip_trsm_workspace_bytes(&need, k, compute);
  if (!work || work_bytes < need)
    return fail(FHIP_E_WORKSPACE, "fhip_trsm_apply: workspace of %zu bytes is smaller than the %zu needed",
                work_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int ik = (int)k;
  if (type == FHIP_C64 && compute == FHIP_C32)
    return launch_wide<double, float, true>((double *)Y, ldy, m, ik, (const float *)work, s);
  if (adjoint) {
    switch (type) {
      case FHIP_C64: return launch_wide<double, double, true, true>((double *)Y, ldy, m, ik, (const double *)work, s);
      case FHIP_C32: return launch_wide<float, float, true, true>((float *)Y, ldy, m, ik, (const float *)work, s);
      case FHIP_F64: return launch_wide<double, double, false, true>((double *)Y, ldy, m, ik, (const double *)work, s);
      case FHIP_F32: return launch_wide<float, float, false, true>((float *)Y, ldy, m, ik, (const float *)work, s);
    }
  }
  switch (type) {
    case FHIP_C64: return launch_wide<double, double, true>((double *)Y, ldy, m, ik, (const double *)work, s);
    case FHIP_C32: return launch_wide<float, float, true>((float *)Y, ldy, m, ik, (const float *)work, s);
    case FHIP_F64: return launch_wide<double, double, false>((double *)Y, ldy, m, ik, (const double *)work, s);
    case FHIP_F32: return launch_wide<float, float, false>((float *)Y, ldy, m, ik, (const float *)work, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_apply: unknown type %d", type);
}
