// Kernel (a): the tall-skinny Gram C <- alpha A^H B + beta C over m rows,
// replacing gram_accumulate! (src/kernels.jl:40, semantics :22-31) as called
// from gram (src/panelblas.jl:33-53) and project (src/panelblas.jl:294).
//
// Design (gfx950):
//  * The output is cut into TILE x TILE tiles (upper-triangle tiles only when
//    A and B are the same matrix) and the m rows into chunks. One workgroup
//    owns one (tile, chunk) pair and sums its rows with v_mfma_f64_16x16x4
//    (or the f32 form for the mixed pass) on the real and imaginary planes,
//    staging KC-row slices of both operands through LDS. A complex product
//    takes three real matrix-core products (3M), not four.
//  * Workgroup ids are dealt round-robin to the 8 XCDs, so the id is split as
//    (xcd, local) and every tile of one chunk runs on the same XCD: the rows
//    of a chunk come from HBM once and the other tiles hit that XCD's L2.
//  * Each workgroup writes its tile to a private partial slot; a second kernel
//    adds the slots in slot order and applies alpha and beta. No atomics, so
//    the bits depend on (m, ka, kb) only.
#include "fhip_internal.hpp"

#include <cmath>
#include <type_traits>

namespace fhip {

template <int TILE>
struct GramShape;
// KC rows are staged per LDS pass; KSPLIT waves split the K steps of one tile
// (the 32 tile), otherwise the 4 waves take the four quadrants of the tile;
// MB x NB is a wave's block count in 16x16 matrix-core tiles; RESIDENT is how
// many such workgroups the MI355X holds at once (256 CUs x LDS/register limit).
#ifndef FHIP_GRAM64_KC
#define FHIP_GRAM64_KC 16
#endif
#ifndef FHIP_GRAM64_RESIDENT
#define FHIP_GRAM64_RESIDENT 1024
#endif
#ifndef FHIP_GRAM_SWIZZLE
#define FHIP_GRAM_SWIZZLE 1
#endif
// Threads of a 64-tile workgroup: 256 (4 waves of 32 x 32 sub-tiles) or 512
// (8 waves of 32 x 16 sub-tiles, 118 registers, four waves a SIMD; it
// measured level with 256 in FP64 and slower in FP32, so 256 is the default).
#ifndef FHIP_GRAM64_THREADS
#define FHIP_GRAM64_THREADS 256
#endif

template <>
struct GramShape<64> {
  static constexpr int KC = FHIP_GRAM64_KC;
  static constexpr int KSPLIT = 1;
  static constexpr int THREADS = FHIP_GRAM64_THREADS;
  static constexpr int MB = 2, NB = THREADS == 256 ? 2 : 1;
  static constexpr int WAVES_EU = THREADS == 256 ? 2 : 4;
  static constexpr int RESIDENT = FHIP_GRAM64_RESIDENT;
};
template <>
struct GramShape<32> {
  static constexpr int KC = 32;
  static constexpr int KSPLIT = 4;
  static constexpr int THREADS = 256;
  static constexpr int MB = 2, NB = 2;
  static constexpr int WAVES_EU = 2;
  static constexpr int RESIDENT = 1024;
};

// Rows staged per LDS pass for the FP32 products of the 64 tile (a tuning
// knob; 32 measured slower than 16 at 629856 x 512).
#ifndef FHIP_GRAM_KC_FP32
#define FHIP_GRAM_KC_FP32 16
#endif
template <typename TC, int TILE>
constexpr int gram_kc() {
  return (sizeof(TC) == 4 && TILE == 64) ? FHIP_GRAM_KC_FP32 : GramShape<TILE>::KC;
}

constexpr int GRAM_MIN_CHUNK_ROWS = 256;

// Upper-triangle tile t -> (ti, tj), tiles numbered column by column.
__host__ __device__ inline void tri_tile(int t, int &ti, int &tj) {
  int j = 0;
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  tj = j;
  ti = t - j * (j + 1) / 2;
}

struct GramPlan {
  int tile;
  int ntm, ntn, ntiles;
  int q;              // chunks = 8 * q
  int64_t chunk_rows; // multiple of KC
  int ksplit;
  int nslots;         // 8 * q * ksplit
};

// The chunk count is chosen so the grid fills whole waves of the resident
// workgroup count of the tile shape, which removes a half-empty last wave. It
// is a constant of the shape, not a device query, so the chunking -- and with
// it every bit of the result -- depends on (m, ka, kb, compute eltype) alone.
constexpr int GRAM_MAX_Q = 256;

static GramPlan gram_plan(int64_t m, int64_t ka, int64_t kb, bool herm, bool fp32) {
  GramPlan p;
  p.tile = (ka > 32 || kb > 32) ? 64 : 32;
  int kc, resident;
  if (p.tile == 64) {
    kc = fp32 ? gram_kc<float, 64>() : gram_kc<double, 64>();
    p.ksplit = GramShape<64>::KSPLIT;
    resident = GramShape<64>::RESIDENT;
  } else {
    kc = GramShape<32>::KC; p.ksplit = GramShape<32>::KSPLIT; resident = GramShape<32>::RESIDENT;
  }
  const int GRAM_RESIDENT_WGS = resident;
  p.ntm = (int)cdiv64(ka, p.tile);
  p.ntn = (int)cdiv64(kb, p.tile);
  p.ntiles = herm ? p.ntm * (p.ntm + 1) / 2 : p.ntm * p.ntn;
  // At most one chunk per GRAM_MIN_CHUNK_ROWS rows.
  int64_t qmax = cdiv64(cdiv64(m > 0 ? m : 1, GRAM_MIN_CHUNK_ROWS), 8);
  if (qmax > GRAM_MAX_Q) qmax = GRAM_MAX_Q;
  int best_q = 1;
  double best = -1.0;
  for (int q = 1; q <= qmax; ++q) {
    const double n = 8.0 * q * p.ntiles;
    const double waves = std::ceil(n / GRAM_RESIDENT_WGS);
    double score = n / (waves * GRAM_RESIDENT_WGS);
    if (n < 2.0 * GRAM_RESIDENT_WGS) score *= n / (2.0 * GRAM_RESIDENT_WGS);
    if (score > best + 1e-9) {
      best = score;
      best_q = q;
    }
    if (n >= 16.0 * GRAM_RESIDENT_WGS) break;
  }
  p.q = best_q;
  int64_t rows = cdiv64(m > 0 ? m : 1, 8 * (int64_t)p.q);
  p.chunk_rows = cdiv64(rows, kc) * kc;
  p.nslots = 8 * p.q * p.ksplit;
  return p;
}

template <typename TIn, typename TC, bool CPLX, int TILE>
__global__ __launch_bounds__(GramShape<TILE>::THREADS) __attribute__((amdgpu_waves_per_eu(GramShape<TILE>::WAVES_EU))) void gram_partial_kernel(
    const TIn *__restrict__ A, int64_t lda, const TIn *__restrict__ B, int64_t ldb, int64_t m, int ka,
    int kb, int64_t chunk_rows, int ntiles, int ntm, int herm, int buf, TC *__restrict__ partial) {
  constexpr int KC = gram_kc<TC, TILE>();
  constexpr int KSPLIT = GramShape<TILE>::KSPLIT;
  constexpr int MB = GramShape<TILE>::MB;
  constexpr int NB = GramShape<TILE>::NB;
  constexpr int GRAM_THREADS = GramShape<TILE>::THREADS;
  // Complex tiles use a swizzled LDS layout with unpadded rows: the 8-byte
  // swizzle for FP64 tiles of 32 or more columns, the 4-byte one for FP32
  // tiles of 64. The others pad each row by 4.
  constexpr int SWZ = !(CPLX && FHIP_GRAM_SWIZZLE) ? 0
                      : (sizeof(TC) == 8 && TILE >= 32) ? 1
                      : (sizeof(TC) == 4 && TILE == 64 && KC == 16) ? 2
                                                                   : 0;
  constexpr int LDS_W = SWZ ? TILE : TILE + 4;
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int PER_THREAD = KC * TILE / GRAM_THREADS;
  __shared__ TC As[NP][KC * LDS_W];
  __shared__ TC Bs[NP][KC * LDS_W];

  const int id = blockIdx.x;
  const int xcd = id & 7;
  const int local = id >> 3;
  const int tile = local % ntiles;
  const int chunk = (local / ntiles) * 8 + xcd;
  int ti, tj;
  if (herm) {
    tri_tile(tile, ti, tj);
  } else {
    ti = tile % ntm;
    tj = tile / ntm;
  }
  const int i0 = ti * TILE;
  const int j0 = tj * TILE;
  const bool same = herm && ti == tj;
  const int64_t r0 = (int64_t)chunk * chunk_rows;
  const int64_t r1 = (r0 + chunk_rows < m) ? r0 + chunk_rows : m;

  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  // The wave's sub-tile (16 MB x 16 NB of the tile, or the whole 32 tile) and
  // the K steps it owns.
  constexpr int WROWS = TILE / (16 * MB);  // sub-tiles down a column of the tile
  static_assert(KSPLIT > 1 || WROWS * (TILE / (16 * NB)) == GRAM_THREADS / 64, "the waves tile the tile");
  // Scalar (wave-uniform) tile offsets: the tests built from them run on the
  // scalar unit.
  const int mi0 = __builtin_amdgcn_readfirstlane((KSPLIT == 1) ? 16 * MB * (w % WROWS) : 0);
  const int nj0 = __builtin_amdgcn_readfirstlane((KSPLIT == 1) ? 16 * NB * (w / WROWS) : 0);
  const int s_begin = (KSPLIT > 1) ? w : 0;
  const int s_stride = KSPLIT;
  // In a diagonal tile of A^H A a sub-tile strictly below the diagonal is
  // never read (the reduce mirrors the upper triangle), so its wave only
  // stages slices.
  const bool skip_quadrant = same && KSPLIT == 1 && mi0 >= nj0 + 16 * NB;

  typedef typename Mfma<TC>::acc acc_t;
  // Complex tiles keep the three 3M sums (p1, p2, p3); real tiles use p1 only.
  acc_t p1[MB][NB], p2[MB][NB], p3[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      p1[a][b] = Mfma<TC>::zero();
      p2[a][b] = Mfma<TC>::zero();
      p3[a][b] = Mfma<TC>::zero();
    }

  // Register staging of the next slice while the current one is multiplied.
  // Unconverted data is loaded straight into the compute registers, and a
  // diagonal tile loads its one operand once. When the stored values are
  // converted (FP32 products of FP64 data) the registers hold the raw stored
  // pairs, converted only when staged, and both operands are loaded on every
  // tile (the second load of a diagonal tile's rows hits the cache): a
  // conversion next to a guarded load, or a branch on the tile kind around
  // the loads, makes the compiler wait for each load on the spot.
  constexpr bool CONVERTED = !std::is_same<TIn, TC>::value;
  typedef typename Pair<TIn>::type P2;
  TC pa_re[PER_THREAD], pa_im[PER_THREAD], pb_re[PER_THREAD], pb_im[PER_THREAD];
  P2 ra[PER_THREAD], rbv[PER_THREAD];
  bool oka[PER_THREAD], okb[PER_THREAD];
  auto fetch = [&](int64_t rb) {
#pragma unroll
    for (int e = 0; e < PER_THREAD; ++e) {
      const int idx = t + GRAM_THREADS * e;
      const int kk = idx % KC;
      const int c = idx / KC;
      const int64_t row = rb + kk;
      const bool rok = row < r1;
      if constexpr (CONVERTED) {
        oka[e] = rok && (i0 + c) < ka;
        ra[e] = load_raw<TIn, CPLX>(A, row + (int64_t)(i0 + c) * lda, oka[e]);
        okb[e] = rok && (j0 + c) < kb;
        rbv[e] = load_raw<TIn, CPLX>(B, row + (int64_t)(j0 + c) * ldb, okb[e]);
      } else {
        load_elem<TIn, TC, CPLX>(A, row + (int64_t)(i0 + c) * lda, rok && (i0 + c) < ka, pa_re[e], pa_im[e]);
        if (!same)
          load_elem<TIn, TC, CPLX>(B, row + (int64_t)(j0 + c) * ldb, rok && (j0 + c) < kb, pb_re[e],
                                   pb_im[e]);
      }
    }
  };
  const bool b_staged = CONVERTED || !same;
  // The fast fetch of a full slice of a tile inside both matrices: a wave's
  // element e covers 64 / KC whole columns of the slice (lane l: row l % KC
  // of column l / KC of its patch), so each is read from a base of its own at
  // the patch (scalar arithmetic) plus the lane's byte offset within a patch,
  // fixed for the whole launch. No row or column test and no 64-bit address
  // arithmetic per element (vector instructions beside the matrix-core
  // products cost matrix-core time on gfx950, about five cycles each). `buf`
  // is the host's word that a patch ((64 / KC - 1) columns and KC rows) lies
  // within a 32-bit byte range; the last, partial slice of a chunk and tiles
  // at the matrices' edges take the guarded fetch above.
  constexpr bool FAST_KIND = CPLX && KSPLIT == 1 && 64 % KC == 0;
  // The fast fetch reads only valid elements, so it needs no range check:
  // FP64 tiles read each patch with global loads from its scalar base (8.75
  // against 8.87 ms per 629,856-row block of the 216^3 basis, and 20 against
  // 26 GB fetched from beyond L2); the FP32 products of FP64 data keep the
  // buffer loads (4.97 against 5.11 ms).
  constexpr bool FAST_GLOBAL = !CONVERTED;
  constexpr int CPW = 64 / KC;  // columns of a wave's patch
  const bool fast = FAST_KIND && buf && i0 + TILE <= ka && j0 + TILE <= kb;
  const int wv = __builtin_amdgcn_readfirstlane(w);
  const uint32_t loff_a = (uint32_t)(((lane % KC) + (int64_t)(lane / KC) * lda) * (int64_t)sizeof(P2));
  const uint32_t loff_b = (uint32_t)(((lane % KC) + (int64_t)(lane / KC) * ldb) * (int64_t)sizeof(P2));
  auto fetch_fast = [&](int64_t rb) {
    if constexpr (FAST_KIND) {
#pragma unroll
      for (int e = 0; e < PER_THREAD; ++e) {
        const int cb = (64 * wv + GRAM_THREADS * e) / KC;
        const TIn *pa = A + 2 * (rb + (int64_t)(i0 + cb) * lda);
        P2 v;
        if constexpr (FAST_GLOBAL)
          v = *reinterpret_cast<const P2 *>(reinterpret_cast<const char *>(pa) + loff_a);
        else
          v = buffer_load_pair<P2>(buffer_rsrc(pa, (uint32_t)(((CPW - 1) * lda + KC) * (int64_t)sizeof(P2))), loff_a);
        if constexpr (CONVERTED) {
          ra[e] = v;
          oka[e] = true;
        } else {
          pa_re[e] = v.x;
          pa_im[e] = v.y;
        }
      }
      if (b_staged) {
#pragma unroll
        for (int e = 0; e < PER_THREAD; ++e) {
          const int cb = (64 * wv + GRAM_THREADS * e) / KC;
          const TIn *pb = B + 2 * (rb + (int64_t)(j0 + cb) * ldb);
          P2 v;
          if constexpr (FAST_GLOBAL)
            v = *reinterpret_cast<const P2 *>(reinterpret_cast<const char *>(pb) + loff_b);
          else
            v = buffer_load_pair<P2>(buffer_rsrc(pb, (uint32_t)(((CPW - 1) * ldb + KC) * (int64_t)sizeof(P2))),
                                     loff_b);
          if constexpr (CONVERTED) {
            rbv[e] = v;
            okb[e] = true;
          } else {
            pb_re[e] = v.x;
            pb_im[e] = v.y;
          }
        }
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int e = 0; e < PER_THREAD; ++e) {
      const int idx = t + GRAM_THREADS * e;
      const int kk = idx % KC;
      const int c = idx / KC;
      const int at = kk * LDS_W + (c ^ swizzle_of<SWZ>(kk));
      if constexpr (CONVERTED) {
        TC re, im;
        from_raw(ra[e], oka[e], re, im);
        As[0][at] = re;
        if (CPLX) As[NP - 1][at] = im;
        from_raw(rbv[e], okb[e], re, im);
        Bs[0][at] = re;
        if (CPLX) Bs[NP - 1][at] = im;
      } else {
        As[0][at] = pa_re[e];
        if (CPLX) As[NP - 1][at] = pa_im[e];
        if (!same) {
          Bs[0][at] = pb_re[e];
          if (CPLX) Bs[NP - 1][at] = pb_im[e];
        }
      }
    }
  };
  auto multiply = [&]() {
    const TC *bre = b_staged ? Bs[0] : As[0];
    const TC *bim = b_staged ? Bs[NP - 1] : As[NP - 1];
    if (skip_quadrant) return;
    if constexpr (CPLX)
      mma_steps_3m<TC, MB, NB, PROD_CONJ_A, SWZ>(As[0], As[NP - 1], LDS_W, mi0, bre, bim, LDS_W, nj0, s_begin,
                                                 KC / 4, s_stride, p1, p2, p3);
    else
      mma_steps<TC, false, MB, NB, PROD_CONJ_A>(As[0], As[0], LDS_W, mi0, bre, bre, LDS_W, nj0, s_begin,
                                                KC / 4, s_stride, p1, p2);
  };
  // The slice loop, in two versions chosen once per workgroup so that each
  // loop's staging registers have one kind of load behind them (loads of
  // both kinds into the same registers make the compiler wait for them where
  // the paths meet, before the products they were to overlap). The fast
  // version runs the chunk's full slices and then its partial last slice, if
  // any, with the guarded fetch.
  auto run = [&](auto fastc) {
    constexpr bool FAST = decltype(fastc)::value;
    const int64_t rend = FAST ? r0 + (r1 - r0) / KC * KC : r1;
    if (r0 < rend) {
      if constexpr (FAST)
        fetch_fast(r0);
      else
        fetch(r0);
    }
    for (int64_t rb = r0; rb < rend; rb += KC) {
      __syncthreads();
      stash();
      __syncthreads();
      if (rb + KC < rend) {
        if constexpr (FAST)
          fetch_fast(rb + KC);
        else
          fetch(rb + KC);
      }
      multiply();
    }
    if constexpr (FAST) {
      if (rend < r1) {
        fetch(rend);
        __syncthreads();
        stash();
        __syncthreads();
        multiply();
      }
    }
  };
  if constexpr (FAST_KIND) {
    if (fast)
      run(std::true_type());
    else
      run(std::false_type());
  } else {
    run(std::false_type());
  }

  // Partial slot of this workgroup (and wave, when waves split K).
  const int slot = chunk * KSPLIT + ((KSPLIT > 1) ? w : 0);
  TC *out = partial + ((int64_t)slot * ntiles + tile) * (TILE * TILE) * NP;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      acc_t re = p1[mb][nb], im = p2[mb][nb];
      if constexpr (CPLX) combine3m<PROD_CONJ_A>(p1[mb][nb], p2[mb][nb], p3[mb][nb], re, im);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = mi0 + 16 * mb + Mfma<TC>::row(lk, r);
        const int jl = nj0 + 16 * nb + li;
        const int64_t off = ((int64_t)jl * TILE + il) * NP;
        out[off] = re[r];
        if (CPLX) out[off + 1] = im[r];
      }
    }
}

// Adds the partial slots and applies alpha and beta. A workgroup takes
// 256 / rg elements; the rg thread groups of an element sum the slots
// s = g, g + rg, ... in order, and an LDS tree adds the groups in a fixed
// pairing. rg is a function of the slot count alone, so the bits depend on
// (m, ka, kb, compute eltype) as the partial sums do.
template <typename TC, typename TOut, bool CPLX, int TILE>
__global__ __launch_bounds__(256) void gram_reduce_kernel(const TC *__restrict__ partial, int nslots,
                                                          int ntiles, int ntm, int herm,
                                                          TOut *__restrict__ C, int64_t ldc, int ka,
                                                          int kb, double al_re, double al_im,
                                                          double be_re, double be_im, int beta_zero, int rg) {
  constexpr int NP = CPLX ? 2 : 1;
  __shared__ double red_re[256], red_im[256];
  const int epb = 256 / rg;
  const int el = threadIdx.x % epb;
  const int g = threadIdx.x / epb;
  const int64_t e = (int64_t)blockIdx.x * epb + el;
  const bool in_grid = e < (int64_t)ntiles * TILE * TILE;
  const int tile = in_grid ? (int)(e / (TILE * TILE)) : 0;
  const int within = in_grid ? (int)(e % (TILE * TILE)) : 0;
  const int jl = within / TILE;
  const int il = within % TILE;
  int ti, tj;
  if (herm) {
    tri_tile(tile, ti, tj);
  } else {
    ti = tile % ntm;
    tj = tile / ntm;
  }
  const int i = ti * TILE + il;
  const int j = tj * TILE + jl;
  // Every thread reaches the barriers below; only live ones load and store.
  const bool live = in_grid && i < ka && j < kb;
  // Below the diagonal of a diagonal tile, take the mirrored upper element so
  // the result is exactly Hermitian.
  const bool lower = herm && ti == tj && il > jl;
  const int src = lower ? (il * TILE + jl) : within;
  double sr = 0.0, si = 0.0;
  const int64_t stride = (int64_t)ntiles * TILE * TILE * NP;
  const TC *p = partial + ((int64_t)g * ntiles * TILE * TILE + (int64_t)tile * TILE * TILE + src) * NP;
  if (live) {
    for (int s = g; s < nslots; s += rg) {
      sr += (double)p[0];
      if (CPLX) si += (double)p[1];
      p += stride * rg;
    }
  }
  red_re[threadIdx.x] = sr;
  red_im[threadIdx.x] = si;
  for (int h = rg / 2; h >= 1; h /= 2) {
    __syncthreads();
    if (g < h) {
      red_re[threadIdx.x] += red_re[threadIdx.x + h * epb];
      red_im[threadIdx.x] += red_im[threadIdx.x + h * epb];
    }
  }
  __syncthreads();
  if (g != 0 || !live) return;
  sr = red_re[threadIdx.x];
  si = red_im[threadIdx.x];
  if (lower) si = -si;
  // The diagonal of A^H A is real.
  if (herm && i == j) si = 0.0;

  auto store = [&](int64_t idx, double vr, double vi) {
    double outr = al_re * vr - al_im * vi;
    double outi = al_re * vi + al_im * vr;
    if (!beta_zero) {
      const double cr = (double)C[idx * NP];
      const double ci = CPLX ? (double)C[idx * NP + 1] : 0.0;
      outr += be_re * cr - be_im * ci;
      outi += be_re * ci + be_im * cr;
    }
    C[idx * NP] = (TOut)outr;
    if (CPLX) C[idx * NP + 1] = (TOut)outi;
  };
  store((int64_t)i + (int64_t)j * ldc, sr, si);
  if (herm && ti < tj && j < ka && i < kb) store((int64_t)j + (int64_t)i * ldc, sr, -si);
}

// Slot groups of the reduce: about 32 slots per thread, a power of two
// between 1 and 256.
static int reduce_groups(int nslots) {
  int rg = 1;
  while (rg < 256 && rg * 64 <= nslots) rg *= 2;
  return rg;
}

template <typename TIn, typename TC, typename TOut, bool CPLX, int TILE>
static int launch_gram(const GramPlan &p, TOut *C, int64_t ldc, const TIn *A, int64_t lda, const TIn *B,
                       int64_t ldb, int64_t m, int ka, int kb, bool herm, double ar, double ai,
                       double br, double bi, void *work, hipStream_t stream) {
  TC *partial = (TC *)work;
  const int grid = 8 * p.q * p.ntiles;
#ifndef FHIP_GRAM_BUF
#define FHIP_GRAM_BUF 1
#endif
  // The fast fetch's descriptors span a wave's patch: 64 / KC columns of KC
  // rows.
  constexpr int KC = gram_kc<TC, TILE>();
  const int64_t span_bytes =
      ((int64_t)(64 / KC - 1) * (lda > ldb ? lda : ldb) + KC) * 2 * (int64_t)sizeof(TIn);
  const int buf = (FHIP_GRAM_BUF && CPLX && span_bytes < (int64_t(1) << 32)) ? 1 : 0;
  hipLaunchKernelGGL((gram_partial_kernel<TIn, TC, CPLX, TILE>), dim3(grid), dim3(GramShape<TILE>::THREADS), 0,
                     stream, A, lda, B, ldb, m, ka, kb, p.chunk_rows, p.ntiles, p.ntm, herm ? 1 : 0, buf,
                     partial);
  int st = check_launch("gram_partial_kernel");
  if (st) return st;
  const int64_t elems = (int64_t)p.ntiles * TILE * TILE;
  const int rg = reduce_groups(p.nslots);
  const int rgrid = (int)cdiv64(elems, 256 / rg);
  const int beta_zero = (br == 0.0 && bi == 0.0) ? 1 : 0;
  hipLaunchKernelGGL((gram_reduce_kernel<TC, TOut, CPLX, TILE>), dim3(rgrid), dim3(256), 0, stream,
                     partial, p.nslots, p.ntiles, p.ntm, herm ? 1 : 0, C, ldc, ka, kb, ar, ai, br, bi,
                     beta_zero, rg);
  return check_launch("gram_reduce_kernel");
}

template <typename TIn, typename TC, typename TOut, bool CPLX>
static int dispatch_gram(const GramPlan &p, void *C, int64_t ldc, const void *A, int64_t lda,
                         const void *B, int64_t ldb, int64_t m, int ka, int kb, bool herm, double ar,
                         double ai, double br, double bi, void *work, hipStream_t stream) {
  if (p.tile == 64)
    return launch_gram<TIn, TC, TOut, CPLX, 64>(p, (TOut *)C, ldc, (const TIn *)A, lda, (const TIn *)B,
                                                ldb, m, ka, kb, herm, ar, ai, br, bi, work, stream);
  return launch_gram<TIn, TC, TOut, CPLX, 32>(p, (TOut *)C, ldc, (const TIn *)A, lda, (const TIn *)B, ldb,
                                              m, ka, kb, herm, ar, ai, br, bi, work, stream);
}

static size_t gram_work_bytes(const GramPlan &p, int compute) {
  const size_t np = is_complex(compute) ? 2 : 1;
  const size_t real = is_double(compute) ? 8 : 4;
  return (size_t)p.nslots * p.ntiles * p.tile * p.tile * np * real;
}

static bool gram_types_ok(int type, int compute) {
  if (!valid_type(type) || !valid_type(compute)) return false;
  if (type == compute) return true;
  return type == FHIP_C64 && compute == FHIP_C32;
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_gram_workspace_bytes(size_t *nbytes, int64_t m, int64_t ka, int64_t kb, int type,
                                         int compute, int hermitian) {
  if (!nbytes) return fail(FHIP_E_ARGUMENT, "fhip_gram_workspace_bytes: null output pointer");
  if (!gram_types_ok(type, compute))
    return fail(FHIP_E_UNSUPPORTED, "fhip_gram: no kernel for storage type %d with compute type %d", type,
                compute);
  if (m < 0 || ka < 0 || kb < 0) return fail(FHIP_E_ARGUMENT, "fhip_gram: negative dimension");
  if (ka == 0 || kb == 0) {
    *nbytes = 0;
    return FHIP_OK;
  }
  GramPlan p = gram_plan(m, ka, kb, hermitian != 0 && ka == kb, !is_double(compute));
  *nbytes = gram_work_bytes(p, compute);
  return FHIP_OK;
}

extern "C" int fhip_gram(void *C, int64_t ldc, const void *A, int64_t lda, const void *B, int64_t ldb,
                         int64_t m, int64_t ka, int64_t kb, double alpha_re, double alpha_im,
                         double beta_re, double beta_im, int type, int compute, void *work,
                         size_t work_bytes, void *stream) {
  if (!gram_types_ok(type, compute))
    return fail(FHIP_E_UNSUPPORTED, "fhip_gram: no kernel for storage type %d with compute type %d", type,
                compute);
  if (m < 0 || ka < 0 || kb < 0) return fail(FHIP_E_ARGUMENT, "fhip_gram: negative dimension");
  if (ka == 0 || kb == 0) return FHIP_OK;
  if (ka > (1 << 20) || kb > (1 << 20)) return fail(FHIP_E_ARGUMENT, "fhip_gram: too many columns");
  if (!C || !A || !B) return fail(FHIP_E_ARGUMENT, "fhip_gram: null matrix pointer");
  if (ldc < ka || (m > 0 && (lda < m || ldb < m)))
    return fail(FHIP_E_ARGUMENT, "fhip_gram: leading dimension smaller than the matrix (ldc=%lld lda=%lld ldb=%lld)",
                (long long)ldc, (long long)lda, (long long)ldb);
  const bool herm = (A == B) && (lda == ldb) && (ka == kb);
  GramPlan p = gram_plan(m, ka, kb, herm, !is_double(compute));
  const size_t need = gram_work_bytes(p, compute);
  if (!work || work_bytes < need)
    return fail(FHIP_E_WORKSPACE, "fhip_gram: workspace of %zu bytes is smaller than the %zu needed",
                work_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int ika = (int)ka, ikb = (int)kb;
  switch (type) {
    case FHIP_C64:
      if (compute == FHIP_C32)
        return dispatch_gram<double, float, double, true>(p, C, ldc, A, lda, B, ldb, m, ika, ikb, herm,
                                                          alpha_re, alpha_im, beta_re, beta_im, work, s);
      return dispatch_gram<double, double, double, true>(p, C, ldc, A, lda, B, ldb, m, ika, ikb, herm,
                                                         alpha_re, alpha_im, beta_re, beta_im, work, s);
    case FHIP_C32:
      return dispatch_gram<float, float, float, true>(p, C, ldc, A, lda, B, ldb, m, ika, ikb, herm, alpha_re,
                                                      alpha_im, beta_re, beta_im, work, s);
    case FHIP_F64:
      return dispatch_gram<double, double, double, false>(p, C, ldc, A, lda, B, ldb, m, ika, ikb, herm,
                                                          alpha_re, alpha_im, beta_re, beta_im, work, s);
    case FHIP_F32:
      return dispatch_gram<float, float, float, false>(p, C, ldc, A, lda, B, ldb, m, ika, ikb, herm,
                                                       alpha_re, alpha_im, beta_re, beta_im, work, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_gram: unknown type %d", type);
}
