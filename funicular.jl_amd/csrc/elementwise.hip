// Elementwise device kernels: eltype conversion (convert_copy!,
// src/kernels.jl:7-20, used by convert_staged! src/panels.jl:365-371 and
// rightmul! src/panelblas.jl:216), a fill that stands in for memset, the
// two matrix-free operators of the benchmark configurations, applied to a
// whole panel (the panel_capable path of apply_operator!, src/operator.jl:80),
// and the steps of the elementwise panel BLAS (scale!, axpy!, dot, norm,
// src/panelblas.jl:303-370).
#include "fhip_internal.hpp"

namespace fhip {

// Copies and fills take a 1-D grid of (column, chunk of ELT_ROWS rows) ids: a
// workgroup's 256 threads each move ELT_ROWS / 256 rows, 256 apart so that a
// wave's accesses are contiguous, with one division per id instead of a
// 64-bit division and remainder per element. One row a thread moves a
// 629,856 x 512 ComplexF64 block at 6.0 TB/s (copy and fill); four rows a
// thread 5.4, the earlier flat grid-stride form 5.0.
#ifndef FHIP_ELT_ROWS
#define FHIP_ELT_ROWS 256
#endif
constexpr int ELT_ROWS = FHIP_ELT_ROWS;
constexpr int ELT_PER_THREAD = ELT_ROWS / 256;

template <typename TD, bool CD, typename TS, bool CS>
__global__ __launch_bounds__(256) void convert_kernel(TD *__restrict__ dst, int64_t ldd,
                                                      const TS *__restrict__ src, int64_t lds, int64_t rows,
                                                      int64_t nchunk, int64_t nids) {
  for (int64_t id = blockIdx.x; id < nids; id += gridDim.x) {
    const int64_t j = id / nchunk;
    const int64_t i0 = (id - j * nchunk) * ELT_ROWS + threadIdx.x;
    TS re[ELT_PER_THREAD], im[ELT_PER_THREAD];
#pragma unroll
    for (int r = 0; r < ELT_PER_THREAD; ++r) {
      const int64_t i = i0 + 256 * r;
      load_elem<TS, TS, CS>(src, (i < rows ? i : 0) + j * lds, true, re[r], im[r]);
    }
#pragma unroll
    for (int r = 0; r < ELT_PER_THREAD; ++r) {
      const int64_t i = i0 + 256 * r;
      if (i < rows) store_elem<TD, CD>(dst, i + j * ldd, (TD)re[r], (TD)im[r]);
    }
  }
}

template <typename T, bool C>
__global__ __launch_bounds__(256) void fill_kernel(T *__restrict__ dst, int64_t ld, int64_t rows, int64_t nchunk,
                                                   int64_t nids, T vr, T vi) {
  for (int64_t id = blockIdx.x; id < nids; id += gridDim.x) {
    const int64_t j = id / nchunk;
    const int64_t i0 = (id - j * nchunk) * ELT_ROWS + threadIdx.x;
#pragma unroll
    for (int r = 0; r < ELT_PER_THREAD; ++r) {
      const int64_t i = i0 + 256 * r;
      if (i < rows) store_elem<T, C>(dst, i + j * ld, vr, vi);
    }
  }
}

// The (column, row chunk) grid of a copy or fill: ids, and workgroups
// launched (each loops over ids when there are more than 2^20).
struct EltGrid {
  int64_t nchunk, nids;
  unsigned blocks;
};
static EltGrid elt_grid(int64_t rows, int64_t cols) {
  EltGrid g;
  g.nchunk = cdiv64(rows, ELT_ROWS);
  g.nids = g.nchunk * cols;
  g.blocks = (unsigned)(g.nids < (int64_t(1) << 20) ? g.nids : (int64_t(1) << 20));
  if (g.blocks < 1) g.blocks = 1;
  return g;
}

static unsigned grid_for(int64_t n) {
  int64_t g = cdiv64(n, 256);
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename TD, bool CD>
static int convert_from(void *dst, int64_t ldd, const void *src, int src_type, int64_t lds, int64_t rows,
                        int64_t cols, hipStream_t s) {
  const EltGrid g = elt_grid(rows, cols);
  switch (src_type) {
    case FHIP_F32:
      hipLaunchKernelGGL((convert_kernel<TD, CD, float, false>), dim3(g.blocks), dim3(256), 0, s, (TD *)dst, ldd,
                         (const float *)src, lds, rows, g.nchunk, g.nids);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((convert_kernel<TD, CD, double, false>), dim3(g.blocks), dim3(256), 0, s, (TD *)dst, ldd,
                         (const double *)src, lds, rows, g.nchunk, g.nids);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((convert_kernel<TD, CD, float, true>), dim3(g.blocks), dim3(256), 0, s, (TD *)dst, ldd,
                         (const float *)src, lds, rows, g.nchunk, g.nids);
      break;
    case FHIP_C64:
      hipLaunchKernelGGL((convert_kernel<TD, CD, double, true>), dim3(g.blocks), dim3(256), 0, s, (TD *)dst, ldd,
                         (const double *)src, lds, rows, g.nchunk, g.nids);
      break;
  }
  return check_launch("convert_kernel");
}

// Periodic tridiagonal stencil, one thread per (row, column) of the panel.
// (Gx)_i = d_i x_i + o (x_{i-1} + x_{i+1}); the adjoint conjugates d and o.
template <typename T>
__global__ __launch_bounds__(256) void tridiag_kernel(T *__restrict__ Y, int64_t ldy, const T *__restrict__ X,
                                                      int64_t ldx, int64_t N, int64_t w, T dr, T di, T orr,
                                                      T oi, const double *__restrict__ diag) {
  const int64_t n = N * w;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int64_t i = e % N;
    const int64_t j = e / N;
    const int64_t im1 = i == 0 ? N - 1 : i - 1;
    const int64_t ip1 = i == N - 1 ? 0 : i + 1;
    T xr, xi, ar, ai, br, bi;
    load_elem<T, T, true>(X, i + j * ldx, true, xr, xi);
    load_elem<T, T, true>(X, im1 + j * ldx, true, ar, ai);
    load_elem<T, T, true>(X, ip1 + j * ldx, true, br, bi);
    T cr = dr, ci = di;
    if (diag) cr += (T)diag[i];
    const T nr = ar + br, ni = ai + bi;
    const T yr = cr * xr - ci * xi + orr * nr - oi * ni;
    const T yi = cr * xi + ci * xr + orr * ni + oi * nr;
    store_elem<T, true>(Y, i + j * ldy, yr, yi);
  }
}

// Periodic 7-point stencil on n1 x n2 x n3, linear index i + n1 (j + n2 l).
// A workgroup sweeps a run of consecutive i, so the +-1 neighbours share
// cache lines and the +-n1, +-n1 n2 neighbours are reused from L2.
template <typename T>
__global__ __launch_bounds__(256) void helmholtz_kernel(T *__restrict__ Y, int64_t ldy,
                                                        const T *__restrict__ X, int64_t ldx, int64_t n1,
                                                        int64_t n2, int64_t n3, int64_t w, T dr, T di, T orr,
                                                        T oi) {
  const int64_t N = n1 * n2 * n3;
  const int64_t n = N * w;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int64_t p = e % N;
    const int64_t col = e / N;
    const int64_t i = p % n1;
    const int64_t jl = p / n1;
    const int64_t j = jl % n2;
    const int64_t l = jl / n2;
    const int64_t base = col * ldx;
    const int64_t s1 = 1, s2 = n1, s3 = n1 * n2;
    const int64_t im = (i == 0 ? p + (n1 - 1) * s1 : p - s1);
    const int64_t ip = (i == n1 - 1 ? p - (n1 - 1) * s1 : p + s1);
    const int64_t jm = (j == 0 ? p + (n2 - 1) * s2 : p - s2);
    const int64_t jp = (j == n2 - 1 ? p - (n2 - 1) * s2 : p + s2);
    const int64_t lm = (l == 0 ? p + (n3 - 1) * s3 : p - s3);
    const int64_t lp = (l == n3 - 1 ? p - (n3 - 1) * s3 : p + s3);
    T xr, xi, sr = 0, si = 0, tr, ti;
    load_elem<T, T, true>(X, base + p, true, xr, xi);
    load_elem<T, T, true>(X, base + im, true, tr, ti);
    sr += tr;
    si += ti;
    load_elem<T, T, true>(X, base + ip, true, tr, ti);
    sr += tr;
    si += ti;
    load_elem<T, T, true>(X, base + jm, true, tr, ti);
    sr += tr;
    si += ti;
    load_elem<T, T, true>(X, base + jp, true, tr, ti);
    sr += tr;
    si += ti;
    load_elem<T, T, true>(X, base + lm, true, tr, ti);
    sr += tr;
    si += ti;
    load_elem<T, T, true>(X, base + lp, true, tr, ti);
    sr += tr;
    si += ti;
    const T yr = dr * xr - di * xi + orr * sr - oi * si;
    const T yi = dr * xi + di * xr + orr * si + oi * sr;
    store_elem<T, true>(Y, p + col * ldy, yr, yi);
  }
}

// The same two stencils on a 1-D grid of (column, 256-row chunk) workgroups,
// for N < 2^31 and fewer than 2^31 workgroups: 32-bit indices, the divisions
// by n1, n2 and the chunk count as multiply-highs, and the workgroups of one
// column renumbered so each XCD sweeps a contiguous range of it. The +-n1 and
// +-n1 n2 neighbours of a row then sit in rows the same XCD read a few hundred
// workgroups earlier, within its L2.
template <typename T>
__global__ __launch_bounds__(256) void tridiag_rows_kernel(T *__restrict__ Y, int64_t ldy, const T *__restrict__ X,
                                                           int64_t ldx, uint32_t N, DivU32 chunks, T dr, T di, T orr,
                                                           T oi, const double *__restrict__ diag) {
  const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t col = divu32(b, chunks);
  const uint32_t i = (b - col * chunks.d) * 256u + threadIdx.x;
  if (i >= N) return;
  const uint32_t im1 = i == 0 ? N - 1 : i - 1;
  const uint32_t ip1 = i == N - 1 ? 0 : i + 1;
  const T *x = X + 2 * (int64_t)col * ldx;
  T xr, xi, ar, ai, br, bi;
  load_elem<T, T, true>(x, i, true, xr, xi);
  load_elem<T, T, true>(x, im1, true, ar, ai);
  load_elem<T, T, true>(x, ip1, true, br, bi);
  T cr = dr, ci = di;
  if (diag) cr += (T)diag[i];
  const T nr = ar + br, ni = ai + bi;
  const T yr = cr * xr - ci * xi + orr * nr - oi * ni;
  const T yi = cr * xi + ci * xr + orr * ni + oi * nr;
  store_elem<T, true>(Y + 2 * (int64_t)col * ldy, i, yr, yi);
}

template <typename T>
__global__ __launch_bounds__(256) void helmholtz_rows_kernel(T *__restrict__ Y, int64_t ldy,
                                                             const T *__restrict__ X, int64_t ldx, uint32_t n1,
                                                             uint32_t n2, uint32_t n3, DivU32 d1, DivU32 d2,
                                                             DivU32 chunks, T dr, T di, T orr, T oi) {
  const uint32_t s3 = n1 * n2, N = s3 * n3;
  const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t col = divu32(b, chunks);
  const uint32_t p = (b - col * chunks.d) * 256u + threadIdx.x;
  if (p >= N) return;
  const uint32_t jl = divu32(p, d1), i = p - jl * n1;
  const uint32_t l = divu32(jl, d2), j = jl - l * n2;
  const uint32_t im = i == 0 ? p + (n1 - 1) : p - 1;
  const uint32_t ip = i == n1 - 1 ? p - (n1 - 1) : p + 1;
  const uint32_t jm = j == 0 ? p + (n2 - 1) * n1 : p - n1;
  const uint32_t jp = j == n2 - 1 ? p - (n2 - 1) * n1 : p + n1;
  const uint32_t lm = l == 0 ? p + (n3 - 1) * s3 : p - s3;
  const uint32_t lp = l == n3 - 1 ? p - (n3 - 1) * s3 : p + s3;
  const T *x = X + 2 * (int64_t)col * ldx;
  T xr, xi, sr, si, tr, ti;
  load_elem<T, T, true>(x, p, true, xr, xi);
  load_elem<T, T, true>(x, im, true, sr, si);
  load_elem<T, T, true>(x, ip, true, tr, ti);
  sr += tr;
  si += ti;
  load_elem<T, T, true>(x, jm, true, tr, ti);
  sr += tr;
  si += ti;
  load_elem<T, T, true>(x, jp, true, tr, ti);
  sr += tr;
  si += ti;
  load_elem<T, T, true>(x, lm, true, tr, ti);
  sr += tr;
  si += ti;
  load_elem<T, T, true>(x, lp, true, tr, ti);
  sr += tr;
  si += ti;
  const T yr = dr * xr - di * xi + orr * sr - oi * si;
  const T yi = dr * xi + di * xr + orr * si + oi * sr;
  store_elem<T, true>(Y + 2 * (int64_t)col * ldy, p, yr, yi);
}

// Shape of the blocked stencil: 256 E points of a plane per workgroup and L
// planes. Two staged planes take 2 (256 E + 2 n1) complex entries of LDS,
// 64 KiB at n1 = 512 in FP64, which is the largest n1 it takes.
// Register sets staging the planes ahead: with two, a plane's loads are
// issued two steps before it is stashed (2.02 against 2.28 ms per 216^3 x 32
// panel with one; three sets drop to two waves a SIMD and take 2.38 ms).
#ifndef FHIP_HELM_STAGES
#define FHIP_HELM_STAGES 2
#endif
#ifndef FHIP_HELM_E
#define FHIP_HELM_E 4
#endif
#ifndef FHIP_HELM_L
#define FHIP_HELM_L 24
#endif
constexpr int HELM_E = FHIP_HELM_E;
constexpr int HELM_L = FHIP_HELM_L;
constexpr int64_t HELM_MAX_N1 = 512;

// 2.5-D blocking of the 7-point stencil. A workgroup owns Q = 256 E
// consecutive points q = i + n1 j of the plane and L consecutive planes l.
// Each plane's points [q0 - n1, q0 + Q + n1) (the owned range and one i-row
// either side, indices taken mod n1 n2, which is the j periodicity) are
// staged in LDS, two planes at a time: the i and j neighbours of a point
// come from its own plane's tile, its l + 1 neighbour from the next plane's
// tile, and its l - 1 neighbour from registers (the centre value kept from
// the plane before). The tile of the plane two ahead is loaded into
// registers FHIP_HELM_STAGES steps before it is stashed, while the planes
// between are computed.
template <typename T, int E, int L>
__global__ __launch_bounds__(256) void helmholtz_tile_kernel(T *__restrict__ Y, int64_t ldy,
                                                             const T *__restrict__ X, int64_t ldx, uint32_t n1,
                                                             uint32_t n2, uint32_t n3, DivU32 d1, DivU32 dq,
                                                             DivU32 dl, T dr, T di, T orr, T oi) {
  typedef typename Pair<T>::type P2;
  extern __shared__ __align__(16) unsigned char helm_lds[];
  constexpr uint32_t Q = 256u * E;
  const uint32_t s3 = n1 * n2;
  const uint32_t span = Q + 2u * n1;  // points per staged plane
  P2 *tile[2] = {reinterpret_cast<P2 *>(helm_lds), reinterpret_cast<P2 *>(helm_lds) + span};
  uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t r1 = divu32(b, dq);
  const uint32_t qt = b - r1 * dq.d;
  const uint32_t col = divu32(r1, dl);
  const uint32_t lg = r1 - col * dl.d;
  const uint32_t q0 = qt * Q;
  const uint32_t l0 = lg * (uint32_t)L;
  const uint32_t nl = n3 - l0 < (uint32_t)L ? n3 - l0 : (uint32_t)L;
  const uint32_t t = threadIdx.x;
  const P2 *x = reinterpret_cast<const P2 *>(X + 2 * (int64_t)col * ldx);
  T *y = Y + 2 * (int64_t)col * ldy;
  // Staging: thread t moves points h = t + 256 r of a plane's span.
  constexpr int R = (int)((256u * E + 2u * 512u + 255u) / 256u);  // enough for n1 <= 512
  const int nr = (int)((span + 255u) / 256u);
  uint32_t gidx[R];
  const uint32_t hbase = q0 + s3 - n1;  // h = 0 is plane point q0 - n1 (mod s3)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    uint32_t g = hbase + t + 256u * (uint32_t)r;
    while (g >= s3) g -= s3;
    gidx[r] = g;
  }
  // Staging register sets: the plane two ahead of the one being computed is
  // loaded NSTG steps before it is stashed.
  constexpr int NSTG = FHIP_HELM_STAGES;
  static_assert(NSTG >= 1 && NSTG <= 3, "one to three staging sets");
  P2 stage[NSTG][R];
  auto fetch = [&](auto set, uint32_t l) {
    constexpr int S = decltype(set)::value;
    const P2 *xp = x + (int64_t)s3 * l;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < nr) stage[S][r] = xp[gidx[r]];
  };
  auto stash = [&](auto set, P2 *dst) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t h = t + 256u * (uint32_t)r;
      if (r < nr && h < span) dst[h] = stage[S][r];
    }
  };
  auto wrap = [&](uint32_t l) {
    while (l >= n3) l -= n3;
    return l;
  };
  // The thread's points and their local (tile) offsets of the i neighbours.
  uint32_t hq[E];
  int dim[E], dip[E];
  bool live[E];
  T pr[E], pi[E];
  const uint32_t lprev = l0 == 0 ? n3 - 1 : l0 - 1;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t q = q0 + 256u * e + t;
    live[e] = q < s3;
    const uint32_t qc = live[e] ? q : s3 - 1;
    const uint32_t j = divu32(qc, d1), i = qc - j * n1;
    hq[e] = q - q0 + n1;
    dim[e] = i == 0 ? (int)n1 - 1 : -1;
    dip[e] = i == n1 - 1 ? 1 - (int)n1 : 1;
    const P2 v = x[(int64_t)s3 * lprev + qc];
    pr[e] = v.x;
    pi[e] = v.y;
  }
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, (NSTG > 1) ? 1 : 0>;
  using S2 = std::integral_constant<int, (NSTG > 2) ? 2 : 0>;
  fetch(S0(), l0);
  stash(S0(), tile[0]);
  fetch(S0(), wrap(l0 + 1));
  stash(S0(), tile[1]);
  if (nl > 1) fetch(S0(), wrap(l0 + 2));
  if constexpr (NSTG >= 2) {
    if (nl > 2) fetch(S1(), wrap(l0 + 3));
  }
  if constexpr (NSTG >= 3) {
    if (nl > 3) fetch(S2(), wrap(l0 + 4));
  }
  __syncthreads();
  // Step s computes plane l0 + s from the tiles of planes l0 + s and
  // l0 + s + 1, then stashes plane l0 + s + 2 over the first and loads plane
  // l0 + s + 2 + NSTG into the set it came from.
  auto step = [&](auto set, uint32_t s) {
    const P2 *cur = tile[s & 1];
    const P2 *nxt = tile[(s & 1) ^ 1];
    const uint32_t l = l0 + s;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t h = hq[e];
      const P2 c = cur[h];
      const P2 a = cur[h + dim[e]];
      const P2 bb = cur[h + dip[e]];
      const P2 jm = cur[h - n1];
      const P2 jp = cur[h + n1];
      const P2 lp = nxt[h];
      T sr = a.x + bb.x, si = a.y + bb.y;
      sr += jm.x;
      si += jm.y;
      sr += jp.x;
      si += jp.y;
      sr += pr[e];
      si += pi[e];
      sr += lp.x;
      si += lp.y;
      const T yr = dr * c.x - di * c.y + orr * sr - oi * si;
      const T yi = dr * c.y + di * c.x + orr * si + oi * sr;
      if (live[e]) store_elem<T, true>(y, (int64_t)s3 * l + (q0 + 256u * e + t), yr, yi);
      pr[e] = c.x;
      pi[e] = c.y;
    }
    if (s + 1 < nl) {
      __syncthreads();
      stash(set, tile[s & 1]);  // plane l + 2 replaces plane l
      if (s + 1 + NSTG < nl) fetch(set, wrap(l + 2 + NSTG));
      __syncthreads();
    }
  };
  if constexpr (NSTG == 3) {
    for (uint32_t s = 0; s < nl; s += 3) {
      step(S0(), s);
      if (s + 1 < nl) step(S1(), s + 1);
      if (s + 2 < nl) step(S2(), s + 2);
    }
  } else if constexpr (NSTG == 2) {
    for (uint32_t s = 0; s < nl; s += 2) {
      step(S0(), s);
      if (s + 1 < nl) step(S1(), s + 1);
    }
  } else {
    for (uint32_t s = 0; s < nl; ++s) step(S0(), s);
  }
}

// Workgroups for the row-chunk stencils, or 0 when the sizes need the 64-bit
// grid-stride forms.
static uint32_t rows_grid(int64_t N, int64_t w, int64_t &chunks) {
  chunks = cdiv64(N, 256);
  if (N >= (int64_t(1) << 31) || chunks * w >= (int64_t(1) << 31)) return 0;
  return (uint32_t)(chunks * w);
}

template <typename T, bool C>
__global__ __launch_bounds__(256) void axpby_kernel(T *__restrict__ Y, int64_t ldy, const T *__restrict__ X,
                                                    int64_t ldx, int64_t rows, int64_t nchunk, int64_t nids, T ar,
                                                    T ai, T br, T bi) {
  for (int64_t id = blockIdx.x; id < nids; id += gridDim.x) {
    const int64_t j = id / nchunk;
    const int64_t i0 = (id - j * nchunk) * ELT_ROWS + threadIdx.x;
#pragma unroll
    for (int r = 0; r < ELT_PER_THREAD; ++r) {
      const int64_t i = i0 + 256 * r;
      if (i >= rows) break;
      T xr, xi, yr, yi;
      load_elem<T, T, C>(X, i + j * ldx, true, xr, xi);
      load_elem<T, T, C>(Y, i + j * ldy, true, yr, yi);
      const T outr = ar * xr - ai * xi + br * yr - bi * yi;
      const T outi = ar * xi + ai * xr + br * yi + bi * yr;
      store_elem<T, C>(Y, i + j * ldy, outr, outi);
    }
  }
}

// out[i] (+)= Re sum_j conj(A[i, j]) B[i, j]: one thread per row, columns
// in order, so the sum is deterministic.
template <typename T, bool C>
__global__ __launch_bounds__(256) void rowdot_kernel(double *__restrict__ out, const T *__restrict__ A,
                                                     int64_t lda, const T *__restrict__ B, int64_t ldb,
                                                     int64_t rows, int64_t cols, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  double s = accumulate ? out[i] : 0.0;
  for (int64_t j = 0; j < cols; ++j) {
    T ar, ai, br, bi;
    load_elem<T, T, C>(A, i + j * lda, true, ar, ai);
    load_elem<T, T, C>(B, i + j * ldb, true, br, bi);
    s += (double)ar * (double)br + (double)ai * (double)bi;
  }
  out[i] = s;
}

// Y <- Y alpha, entry by entry, as Julia's rmul! computes it
// (src/panelblas.jl:311-316): a real scalar scales both parts, a complex one
// is the schoolbook product. Contraction is off so each part is one rounded
// product (and one rounded sum for a complex scalar), as on the CPU.
template <typename T, bool C>
__global__ __launch_bounds__(256) void scale_kernel(T *__restrict__ Y, int64_t ldy, int64_t rows, T ar, T ai,
                                                    int real_scalar) {
#pragma clang fp contract(off)
  const int64_t j = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < rows; i += (int64_t)gridDim.x * 256) {
    T yr, yi;
    load_elem<T, T, C>(Y, i + j * ldy, true, yr, yi);
    T outr, outi;
    if (!C || real_scalar) {
      outr = yr * ar;
      outi = yi * ar;
    } else {
      outr = yr * ar - yi * ai;
      outi = yr * ai + yi * ar;
    }
    store_elem<T, C>(Y, i + j * ldy, outr, outi);
  }
}

// y <- a y + b entry by entry, the product formed as scale_kernel forms it
// and the constant added after it, each part rounded once per operation.
template <typename T, bool C>
__global__ __launch_bounds__(256) void affine_kernel(T *__restrict__ Y, int64_t ldy, int64_t rows, T ar, T ai,
                                                     int real_scalar, T br, T bi) {
#pragma clang fp contract(off)
  const int64_t j = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < rows; i += (int64_t)gridDim.x * 256) {
    T yr, yi;
    load_elem<T, T, C>(Y, i + j * ldy, true, yr, yi);
    T pr, pi;
    if (!C || real_scalar) {
      pr = yr * ar;
      pi = yi * ar;
    } else {
      pr = yr * ar - yi * ai;
      pi = yr * ai + yi * ar;
    }
    store_elem<T, C>(Y, i + j * ldy, pr + br, pi + bi);
  }
}

// Sum over a rows x cols block of conj(x) y (mode 0) or |x|^2 (mode 1), in
// FP64. Workgroup (chunk, column) sums DOT_ROWS rows of one column, each
// thread DOT_PER_THREAD of them at a stride of 256, then an LDS tree in a
// fixed pairing; the partials go to slot column * nchunks + chunk. The order
// of every addition depends only on (rows, cols).
constexpr int DOT_THREADS = 256;
constexpr int DOT_PER_THREAD = 16;
constexpr int64_t DOT_ROWS = (int64_t)DOT_THREADS * DOT_PER_THREAD;

template <int N>
__device__ inline void tree_sum(double (&sh)[N][DOT_THREADS], double (&v)[N]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < N; ++q) sh[q][t] = v[q];
  __syncthreads();
  for (int half = DOT_THREADS / 2; half > 0; half >>= 1) {
    if (t < half) {
#pragma unroll
      for (int q = 0; q < N; ++q) sh[q][t] += sh[q][t + half];
    }
    __syncthreads();
  }
}

template <typename T, bool C>
__global__ __launch_bounds__(DOT_THREADS) void dot_partial_kernel(double *__restrict__ partial,
                                                                  const T *__restrict__ X, int64_t ldx,
                                                                  const T *__restrict__ Y, int64_t ldy,
                                                                  int64_t rows, int64_t nchunks, int mode) {
  const int64_t chunk = blockIdx.x;
  const int64_t j = blockIdx.y;
  const int64_t r0 = chunk * DOT_ROWS;
  double sre = 0.0, sim = 0.0;
#pragma unroll 4
  for (int e = 0; e < DOT_PER_THREAD; ++e) {
    const int64_t i = r0 + (int64_t)DOT_THREADS * e + threadIdx.x;
    if (i < rows) {
      T xr, xi;
      load_elem<T, T, C>(X, i + j * ldx, true, xr, xi);
      if (mode == 1) {
        sre += (double)xr * (double)xr + (double)xi * (double)xi;
      } else {
        T yr, yi;
        load_elem<T, T, C>(Y, i + j * ldy, true, yr, yi);
        sre += (double)xr * (double)yr + (double)xi * (double)yi;
        sim += (double)xr * (double)yi - (double)xi * (double)yr;
      }
    }
  }
  __shared__ double sh[2][DOT_THREADS];
  double v[2] = {sre, sim};
  tree_sum<2>(sh, v);
  if (threadIdx.x == 0) {
    const int64_t slot = j * nchunks + chunk;
    partial[2 * slot] = sh[0][0];
    partial[2 * slot + 1] = sh[1][0];
  }
}

// out <- (accumulate ? out : 0) + the sum of n partials, one workgroup:
// thread t adds slots t, t + 256, ... in order, then the LDS tree.
__global__ __launch_bounds__(DOT_THREADS) void dot_finish_kernel(double *__restrict__ out,
                                                                 const double *__restrict__ partial, int64_t n,
                                                                 int accumulate) {
  double sre = 0.0, sim = 0.0;
  for (int64_t q = threadIdx.x; q < n; q += DOT_THREADS) {
    sre += partial[2 * q];
    sim += partial[2 * q + 1];
  }
  __shared__ double sh[2][DOT_THREADS];
  double v[2] = {sre, sim};
  tree_sum<2>(sh, v);
  if (threadIdx.x == 0) {
    out[0] = (accumulate ? out[0] : 0.0) + sh[0][0];
    out[1] = (accumulate ? out[1] : 0.0) + sh[1][0];
  }
}

static int64_t dot_chunks(int64_t rows) { return cdiv64(rows, DOT_ROWS); }

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_rowdot(double *out, const void *A, int64_t lda, const void *B, int64_t ldb, int64_t rows,
                           int64_t cols, int type, int accumulate, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_rowdot: unknown eltype");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_rowdot: negative dimension");
  if (rows == 0) return FHIP_OK;
  if (!out || (cols > 0 && (!A || !B))) return fail(FHIP_E_ARGUMENT, "fhip_rowdot: null pointer");
  if (lda < rows || ldb < rows) return fail(FHIP_E_ARGUMENT, "fhip_rowdot: leading dimension smaller than the rows");
  const unsigned g = (unsigned)cdiv64(rows, 256);
  hipStream_t s = (hipStream_t)stream;
  switch (type) {
    case FHIP_F32:
      hipLaunchKernelGGL((rowdot_kernel<float, false>), dim3(g), dim3(256), 0, s, out, (const float *)A, lda,
                         (const float *)B, ldb, rows, cols, accumulate);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((rowdot_kernel<double, false>), dim3(g), dim3(256), 0, s, out, (const double *)A, lda,
                         (const double *)B, ldb, rows, cols, accumulate);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((rowdot_kernel<float, true>), dim3(g), dim3(256), 0, s, out, (const float *)A, lda,
                         (const float *)B, ldb, rows, cols, accumulate);
      break;
    case FHIP_C64:
      hipLaunchKernelGGL((rowdot_kernel<double, true>), dim3(g), dim3(256), 0, s, out, (const double *)A, lda,
                         (const double *)B, ldb, rows, cols, accumulate);
      break;
  }
  return check_launch("rowdot_kernel");
}

extern "C" int fhip_axpby(void *Y, int64_t ldy, const void *X, int64_t ldx, int64_t rows, int64_t cols,
                          double a_re, double a_im, double b_re, double b_im, int type, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_axpby: unknown eltype");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_axpby: negative dimension");
  if (rows == 0 || cols == 0) return FHIP_OK;
  if (!Y || !X) return fail(FHIP_E_ARGUMENT, "fhip_axpby: null pointer");
  if (ldy < rows || ldx < rows) return fail(FHIP_E_ARGUMENT, "fhip_axpby: leading dimension smaller than the rows");
  const EltGrid g = elt_grid(rows, cols);
  hipStream_t s = (hipStream_t)stream;
  switch (type) {
    case FHIP_F32:
      hipLaunchKernelGGL((axpby_kernel<float, false>), dim3(g.blocks), dim3(256), 0, s, (float *)Y, ldy,
                         (const float *)X, ldx, rows, g.nchunk, g.nids, (float)a_re, 0.0f, (float)b_re, 0.0f);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((axpby_kernel<double, false>), dim3(g.blocks), dim3(256), 0, s, (double *)Y, ldy,
                         (const double *)X, ldx, rows, g.nchunk, g.nids, a_re, 0.0, b_re, 0.0);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((axpby_kernel<float, true>), dim3(g.blocks), dim3(256), 0, s, (float *)Y, ldy,
                         (const float *)X, ldx, rows, g.nchunk, g.nids, (float)a_re, (float)a_im, (float)b_re,
                         (float)b_im);
      break;
    case FHIP_C64:
      hipLaunchKernelGGL((axpby_kernel<double, true>), dim3(g.blocks), dim3(256), 0, s, (double *)Y, ldy,
                         (const double *)X, ldx, rows, g.nchunk, g.nids, a_re, a_im, b_re, b_im);
      break;
  }
  return check_launch("axpby_kernel");
}

extern "C" int fhip_convert_copy(void *dst, int dst_type, int64_t ldd, const void *src, int src_type,
                                 int64_t lds, int64_t rows, int64_t cols, void *stream) {
  if (!valid_type(dst_type) || !valid_type(src_type))
    return fail(FHIP_E_UNSUPPORTED, "fhip_convert_copy: unknown eltype");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_convert_copy: negative dimension");
  if (rows == 0 || cols == 0) return FHIP_OK;
  if (!dst || !src) return fail(FHIP_E_ARGUMENT, "fhip_convert_copy: null pointer");
  if (ldd < rows || lds < rows) return fail(FHIP_E_ARGUMENT, "fhip_convert_copy: leading dimension smaller than the rows");
  if (is_complex(src_type) && !is_complex(dst_type))
    return fail(FHIP_E_ARGUMENT, "fhip_convert_copy: a complex source cannot be converted to a real destination");
  hipStream_t s = (hipStream_t)stream;
  switch (dst_type) {
    case FHIP_F32: return convert_from<float, false>(dst, ldd, src, src_type, lds, rows, cols, s);
    case FHIP_F64: return convert_from<double, false>(dst, ldd, src, src_type, lds, rows, cols, s);
    case FHIP_C32: return convert_from<float, true>(dst, ldd, src, src_type, lds, rows, cols, s);
    case FHIP_C64: return convert_from<double, true>(dst, ldd, src, src_type, lds, rows, cols, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_convert_copy: unknown eltype");
}

extern "C" int fhip_fill(void *dst, int type, int64_t ld, int64_t rows, int64_t cols, double value_re,
                         double value_im, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_fill: unknown eltype");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_fill: negative dimension");
  if (rows == 0 || cols == 0) return FHIP_OK;
  if (!dst) return fail(FHIP_E_ARGUMENT, "fhip_fill: null pointer");
  if (ld < rows) return fail(FHIP_E_ARGUMENT, "fhip_fill: leading dimension smaller than the rows");
  const EltGrid g = elt_grid(rows, cols);
  hipStream_t s = (hipStream_t)stream;
  switch (type) {
    case FHIP_F32:
      hipLaunchKernelGGL((fill_kernel<float, false>), dim3(g.blocks), dim3(256), 0, s, (float *)dst, ld, rows,
                         g.nchunk, g.nids, (float)value_re, 0.0f);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((fill_kernel<double, false>), dim3(g.blocks), dim3(256), 0, s, (double *)dst, ld, rows,
                         g.nchunk, g.nids, value_re, 0.0);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((fill_kernel<float, true>), dim3(g.blocks), dim3(256), 0, s, (float *)dst, ld, rows,
                         g.nchunk, g.nids, (float)value_re, (float)value_im);
      break;
    case FHIP_C64:
      hipLaunchKernelGGL((fill_kernel<double, true>), dim3(g.blocks), dim3(256), 0, s, (double *)dst, ld, rows,
                         g.nchunk, g.nids, value_re, value_im);
      break;
  }
  return check_launch("fill_kernel");
}

extern "C" int fhip_stencil_tridiag(void *Y, int64_t ldy, const void *X, int64_t ldx, int64_t N, int64_t w,
                                    double d_re, double d_im, double o_re, double o_im, const double *diag,
                                    int adjoint, int type, void *stream) {
  if (type != FHIP_C64 && type != FHIP_C32)
    return fail(FHIP_E_UNSUPPORTED, "fhip_stencil_tridiag: the stencil operators are complex");
  if (N < 0 || w < 0) return fail(FHIP_E_ARGUMENT, "fhip_stencil_tridiag: negative dimension");
  if (N == 0 || w == 0) return FHIP_OK;
  if (!Y || !X) return fail(FHIP_E_ARGUMENT, "fhip_stencil_tridiag: null pointer");
  if (X == Y) return fail(FHIP_E_ARGUMENT, "fhip_stencil_tridiag: output aliases input");
  if (ldx < N || ldy < N) return fail(FHIP_E_ARGUMENT, "fhip_stencil_tridiag: leading dimension smaller than N");
  if (adjoint) {
    d_im = -d_im;
    o_im = -o_im;
  }
  hipStream_t s = (hipStream_t)stream;
  int64_t chunks;
  if (const uint32_t nb = rows_grid(N, w, chunks)) {
    const DivU32 dc = make_divu32((uint32_t)chunks);
    if (type == FHIP_C64)
      hipLaunchKernelGGL((tridiag_rows_kernel<double>), dim3(nb), dim3(256), 0, s, (double *)Y, ldy,
                         (const double *)X, ldx, (uint32_t)N, dc, d_re, d_im, o_re, o_im, diag);
    else
      hipLaunchKernelGGL((tridiag_rows_kernel<float>), dim3(nb), dim3(256), 0, s, (float *)Y, ldy,
                         (const float *)X, ldx, (uint32_t)N, dc, (float)d_re, (float)d_im, (float)o_re,
                         (float)o_im, diag);
    return check_launch("tridiag_rows_kernel");
  }
  const unsigned g = grid_for(N * w);
  if (type == FHIP_C64)
    hipLaunchKernelGGL((tridiag_kernel<double>), dim3(g), dim3(256), 0, s, (double *)Y, ldy, (const double *)X,
                       ldx, N, w, d_re, d_im, o_re, o_im, diag);
  else
    hipLaunchKernelGGL((tridiag_kernel<float>), dim3(g), dim3(256), 0, s, (float *)Y, ldy, (const float *)X, ldx,
                       N, w, (float)d_re, (float)d_im, (float)o_re, (float)o_im, diag);
  return check_launch("tridiag_kernel");
}

extern "C" int fhip_stencil_helmholtz3d(void *Y, int64_t ldy, const void *X, int64_t ldx, int64_t n1,
                                        int64_t n2, int64_t n3, int64_t w, double d_re, double d_im,
                                        double o_re, double o_im, int adjoint, int type, void *stream) {
  if (type != FHIP_C64 && type != FHIP_C32)
    return fail(FHIP_E_UNSUPPORTED, "fhip_stencil_helmholtz3d: the stencil operators are complex");
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || w < 0)
    return fail(FHIP_E_ARGUMENT, "fhip_stencil_helmholtz3d: grid dimensions must be positive");
  if (w == 0) return FHIP_OK;
  const int64_t N = n1 * n2 * n3;
  if (!Y || !X) return fail(FHIP_E_ARGUMENT, "fhip_stencil_helmholtz3d: null pointer");
  if (X == Y) return fail(FHIP_E_ARGUMENT, "fhip_stencil_helmholtz3d: output aliases input");
  if (ldx < N || ldy < N) return fail(FHIP_E_ARGUMENT, "fhip_stencil_helmholtz3d: leading dimension smaller than N");
  if (adjoint) {
    d_im = -d_im;
    o_im = -o_im;
  }
  hipStream_t s = (hipStream_t)stream;
  int64_t chunks;
  const int64_t s3 = n1 * n2;
  const int64_t qtiles = cdiv64(s3, 256 * HELM_E), lgroups = cdiv64(n3, HELM_L);
  if (N < (int64_t(1) << 31) && s3 < (int64_t(1) << 30) && n1 <= HELM_MAX_N1 &&
      qtiles * lgroups * w < (int64_t(1) << 31)) {
    const DivU32 d1 = make_divu32((uint32_t)n1), dq = make_divu32((uint32_t)qtiles);
    const DivU32 dl = make_divu32((uint32_t)lgroups);
    const unsigned nb = (unsigned)(qtiles * lgroups * w);
    const size_t span = 256 * HELM_E + 2 * n1;
    if (type == FHIP_C64)
      hipLaunchKernelGGL((helmholtz_tile_kernel<double, HELM_E, HELM_L>), dim3(nb), dim3(256), 2 * span * 16, s,
                         (double *)Y, ldy, (const double *)X, ldx, (uint32_t)n1, (uint32_t)n2, (uint32_t)n3, d1,
                         dq, dl, d_re, d_im, o_re, o_im);
    else
      hipLaunchKernelGGL((helmholtz_tile_kernel<float, HELM_E, HELM_L>), dim3(nb), dim3(256), 2 * span * 8, s,
                         (float *)Y, ldy, (const float *)X, ldx, (uint32_t)n1, (uint32_t)n2, (uint32_t)n3, d1, dq,
                         dl, (float)d_re, (float)d_im, (float)o_re, (float)o_im);
    return check_launch("helmholtz_tile_kernel");
  }
  if (const uint32_t nb = rows_grid(N, w, chunks)) {
    const DivU32 d1 = make_divu32((uint32_t)n1), d2 = make_divu32((uint32_t)n2);
    const DivU32 dc = make_divu32((uint32_t)chunks);
    if (type == FHIP_C64)
      hipLaunchKernelGGL((helmholtz_rows_kernel<double>), dim3(nb), dim3(256), 0, s, (double *)Y, ldy,
                         (const double *)X, ldx, (uint32_t)n1, (uint32_t)n2, (uint32_t)n3, d1, d2, dc, d_re,
                         d_im, o_re, o_im);
    else
      hipLaunchKernelGGL((helmholtz_rows_kernel<float>), dim3(nb), dim3(256), 0, s, (float *)Y, ldy,
                         (const float *)X, ldx, (uint32_t)n1, (uint32_t)n2, (uint32_t)n3, d1, d2, dc,
                         (float)d_re, (float)d_im, (float)o_re, (float)o_im);
    return check_launch("helmholtz_rows_kernel");
  }
  const unsigned g = grid_for(N * w);
  if (type == FHIP_C64)
    hipLaunchKernelGGL((helmholtz_kernel<double>), dim3(g), dim3(256), 0, s, (double *)Y, ldy,
                       (const double *)X, ldx, n1, n2, n3, w, d_re, d_im, o_re, o_im);
  else
    hipLaunchKernelGGL((helmholtz_kernel<float>), dim3(g), dim3(256), 0, s, (float *)Y, ldy, (const float *)X,
                       ldx, n1, n2, n3, w, (float)d_re, (float)d_im, (float)o_re, (float)o_im);
  return check_launch("helmholtz_kernel");
}

extern "C" int fhip_scale(void *Y, int64_t ldy, int64_t rows, int64_t cols, double a_re, double a_im,
                          int real_scalar, int type, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_scale: unknown eltype");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_scale: negative dimension");
  if (rows == 0 || cols == 0) return FHIP_OK;
  if (!Y) return fail(FHIP_E_ARGUMENT, "fhip_scale: null pointer");
  if (ldy < rows) return fail(FHIP_E_ARGUMENT, "fhip_scale: leading dimension smaller than the rows");
  if (cols > 65535) return fail(FHIP_E_ARGUMENT, "fhip_scale: more than 65535 columns in one call");
  int64_t gx = cdiv64(rows, 256);
  if (gx > 4096) gx = 4096;
  const dim3 grid((unsigned)gx, (unsigned)cols);
  hipStream_t s = (hipStream_t)stream;
  switch (type) {
    case FHIP_F32:
      hipLaunchKernelGGL((scale_kernel<float, false>), grid, dim3(256), 0, s, (float *)Y, ldy, rows, (float)a_re,
                         0.0f, 1);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((scale_kernel<double, false>), grid, dim3(256), 0, s, (double *)Y, ldy, rows, a_re, 0.0, 1);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((scale_kernel<float, true>), grid, dim3(256), 0, s, (float *)Y, ldy, rows, (float)a_re,
                         (float)a_im, real_scalar);
      break;
    case FHIP_C64:
      hipLaunchKernelGGL((scale_kernel<double, true>), grid, dim3(256), 0, s, (double *)Y, ldy, rows, a_re, a_im,
                         real_scalar);
      break;
  }
  return check_launch("scale_kernel");
}

extern "C" int fhip_affine(void *Y, int64_t ldy, int64_t rows, int64_t cols, double a_re, double a_im,
                           int real_scalar, double b_re, double b_im, int type, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_affine: unknown eltype");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_affine: negative dimension");
  if (rows == 0 || cols == 0) return FHIP_OK;
  if (!Y) return fail(FHIP_E_ARGUMENT, "fhip_affine: null pointer");
  if (ldy < rows) return fail(FHIP_E_ARGUMENT, "fhip_affine: leading dimension smaller than the rows");
  if (cols > 65535) return fail(FHIP_E_ARGUMENT, "fhip_affine: more than 65535 columns in one call");
  int64_t gx = cdiv64(rows, 256);
  if (gx > 4096) gx = 4096;
  const dim3 grid((unsigned)gx, (unsigned)cols);
  hipStream_t s = (hipStream_t)stream;
  switch (type) {
    case FHIP_F32:
      hipLaunchKernelGGL((affine_kernel<float, false>), grid, dim3(256), 0, s, (float *)Y, ldy, rows, (float)a_re,
                         0.0f, 1, (float)b_re, 0.0f);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((affine_kernel<double, false>), grid, dim3(256), 0, s, (double *)Y, ldy, rows, a_re, 0.0,
                         1, b_re, 0.0);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((affine_kernel<float, true>), grid, dim3(256), 0, s, (float *)Y, ldy, rows, (float)a_re,
                         (float)a_im, real_scalar, (float)b_re, (float)b_im);
      break;
    case FHIP_C64:
      hipLaunchKernelGGL((affine_kernel<double, true>), grid, dim3(256), 0, s, (double *)Y, ldy, rows, a_re, a_im,
                         real_scalar, b_re, b_im);
      break;
  }
  return check_launch("affine_kernel");
}

extern "C" int fhip_dot_workspace_bytes(size_t *nbytes, int64_t rows, int64_t cols) {
  if (!nbytes) return fail(FHIP_E_ARGUMENT, "fhip_dot_workspace_bytes: null output pointer");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_dot_workspace_bytes: negative dimension");
  *nbytes = (size_t)(dot_chunks(rows) * cols) * 2 * sizeof(double);
  return FHIP_OK;
}

extern "C" int fhip_dot(double *out, const void *X, int64_t ldx, const void *Y, int64_t ldy, int64_t rows,
                        int64_t cols, int type, int mode, int accumulate, void *work, size_t work_bytes,
                        void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_dot: unknown eltype");
  if (mode != 0 && mode != 1) return fail(FHIP_E_ARGUMENT, "fhip_dot: mode must be 0 (conj(x) y) or 1 (|x|^2)");
  if (rows < 0 || cols < 0) return fail(FHIP_E_ARGUMENT, "fhip_dot: negative dimension");
  if (!out) return fail(FHIP_E_ARGUMENT, "fhip_dot: null output pointer");
  if (cols > 65535) return fail(FHIP_E_ARGUMENT, "fhip_dot: more than 65535 columns in one call");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nchunks = dot_chunks(rows);
  const int64_t nslots = nchunks * cols;
  if (nslots > 0) {
    if (!X || (mode == 0 && !Y)) return fail(FHIP_E_ARGUMENT, "fhip_dot: null matrix pointer");
    if (ldx < rows || (mode == 0 && ldy < rows))
      return fail(FHIP_E_ARGUMENT, "fhip_dot: leading dimension smaller than the rows");
    if (nchunks > 0x7fffffff) return fail(FHIP_E_ARGUMENT, "fhip_dot: too many rows");
    const size_t need = (size_t)nslots * 2 * sizeof(double);
    if (!work || work_bytes < need)
      return fail(FHIP_E_ARGUMENT, "fhip_dot: workspace of %zu bytes, %zu needed", work_bytes, need);
    const dim3 grid((unsigned)nchunks, (unsigned)cols);
    double *partial = (double *)work;
    switch (type) {
      case FHIP_F32:
        hipLaunchKernelGGL((dot_partial_kernel<float, false>), grid, dim3(DOT_THREADS), 0, s, partial,
                           (const float *)X, ldx, (const float *)Y, ldy, rows, nchunks, mode);
        break;
      case FHIP_F64:
        hipLaunchKernelGGL((dot_partial_kernel<double, false>), grid, dim3(DOT_THREADS), 0, s, partial,
                           (const double *)X, ldx, (const double *)Y, ldy, rows, nchunks, mode);
        break;
      case FHIP_C32:
        hipLaunchKernelGGL((dot_partial_kernel<float, true>), grid, dim3(DOT_THREADS), 0, s, partial,
                           (const float *)X, ldx, (const float *)Y, ldy, rows, nchunks, mode);
        break;
      case FHIP_C64:
        hipLaunchKernelGGL((dot_partial_kernel<double, true>), grid, dim3(DOT_THREADS), 0, s, partial,
                           (const double *)X, ldx, (const double *)Y, ldy, rows, nchunks, mode);
        break;
    }
    const int st = check_launch("dot_partial_kernel");
    if (st) return st;
  }
  hipLaunchKernelGGL(dot_finish_kernel, dim3(1), dim3(DOT_THREADS), 0, s, out, (const double *)work, nslots,
                     accumulate);
  return check_launch("dot_finish_kernel");
}
