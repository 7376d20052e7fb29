// Kernel (d): the seeded GhostPanels fill, replacing generate_panel! and
// stage_host!(::GhostTier) (src/ghosts.jl:97-122), which fill a column with
// randn! from Xoshiro(hash(col, hash(seed))) on the host producer task.
//
// The reference's stream is Julia's hash + Xoshiro256++ + ziggurat, exact only
// within one Julia version (src/ghosts.jl:34-38). This fill keeps its
// contract -- column col depends on (seed, col) alone, so panels regenerate
// in any order and at any width, entries are standard normal, complex parts
// have variance 1/2 -- and makes it parallel over rows: entry draws are
// counter based (a SplitMix64 output per counter), so any (row, column) is
// generated without the rows before it.
//
// A normal comes from a 256-layer ziggurat (Marsaglia and Tsang, 2000). Its
// fast path is one integer compare and one multiply. The wedge and tail tests
// need exp and log; they use the fixed operation sequences det_exp/det_log
// below (adds, multiplies and fused multiply-adds only, contraction off), so
// the host and the device produce the same bits for every draw.
#include "fhip_internal.hpp"

#include <cmath>
#include <mutex>

#pragma clang fp contract(off)

namespace fhip {

#define GH_GAMMA 0x9E3779B97F4A7C15ull
#define GH_SEEDK 0x6A09E667F3BCC909ull
#define GH_LANE1 0xB7E151628AED2A6Bull
#define GH_LANE2 0x243F6A8885A308D3ull
#define GH_R 3.6541528853610088
#define GH_V 4.92867323399e-3

__host__ __device__ inline uint64_t gh_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t gh_colkey(uint64_t seed, int64_t col) {
  return gh_mix(gh_mix(seed ^ GH_SEEDK) + GH_GAMMA * (uint64_t)col);
}

__host__ __device__ inline uint64_t gh_word(uint64_t key, uint64_t lanek, uint64_t c) {
  return gh_mix((key ^ lanek) + GH_GAMMA * (c + 1));
}

__host__ __device__ inline double gh_bits_to_double(uint64_t b) {
  union {
    uint64_t u;
    double d;
  } v;
  v.u = b;
  return v.d;
}

__host__ __device__ inline uint64_t gh_double_to_bits(double d) {
  union {
    uint64_t u;
    double d;
  } v;
  v.d = d;
  return v.u;
}

// exp(x) for x <= 0: Cody-Waite reduction by ln 2 and a degree-13 Taylor
// polynomial in Horner form.
__host__ __device__ inline double det_exp(double x) {
  if (x < -745.0) return 0.0;
  const double t = x * 1.4426950408889634;
  const double magic = 6755399441055744.0;
  const double n = (t + magic) - magic;
  double r = fma(-n, 6.93147180369123816490e-01, x);
  r = fma(-n, 1.90821492927058770002e-10, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  int64_t k = (int64_t)n;
  if (k < -1000) {
    p = p * gh_bits_to_double((uint64_t)(1023 - 600) << 52);  // 2^-600
    k += 600;
  }
  return p * gh_bits_to_double((uint64_t)(k + 1023) << 52);
}

// log(x) for a normal x > 0: x = 2^e m with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1)/(m + 1), the reciprocal by Newton steps.
__host__ __device__ inline double det_log(double x) {
  const uint64_t b = gh_double_to_bits(x);
  int64_t e = (int64_t)((b >> 52) & 0x7FF) - 1023;
  double m = gh_bits_to_double((b & 0xFFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double d = 2.0 + f;
  double y = 0.5;
  for (int i = 0; i < 6; ++i) {
    const double err = fma(-d, y, 1.0);
    y = fma(y, err, y);
  }
  const double s = f * y;
  const double z = s * s;
  double q = 1.0 / 23.0;
  q = fma(q, z, 1.0 / 21.0);
  q = fma(q, z, 1.0 / 19.0);
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  q = fma(q, z, 1.0);
  const double lnm = 2.0 * s * q;
  const double ed = (double)e;
  return fma(ed, 6.93147180369123816490e-01, fma(ed, 1.90821492927058770002e-10, lnm));
}

struct ZigTables {
  double wt[256];      // layer edge x_i scaled by 2^-53
  double f[257];       // exp(-x_i^2 / 2), f[256] = 1
  uint64_t kt[256];    // fast-accept bound on the 53-bit uniform
};

static void build_tables(ZigTables &z) {
  double x[257];
  auto dens = [](double v) { return det_exp(-0.5 * (v * v)); };
  x[1] = GH_R;
  x[0] = GH_V / dens(GH_R);
  for (int i = 1; i < 255; ++i) {
    const double a = GH_V / x[i] + dens(x[i]);
    x[i + 1] = a >= 1.0 ? 0.0 : std::sqrt(-2.0 * det_log(a));
  }
  x[256] = 0.0;
  const double two53 = 9007199254740992.0;
  for (int i = 0; i < 256; ++i) {
    z.wt[i] = x[i] / two53;
    z.f[i] = dens(x[i]);
    const double ratio = (i == 0 ? GH_R : x[i + 1]) / x[i];
    z.kt[i] = (uint64_t)(ratio * two53);
  }
  z.f[256] = 1.0;
}

__constant__ ZigTables g_zig;

// The draw-th normal of a column's stream, reading the ziggurat tables from
// wt, f and kt (the constant-memory copy, or a workgroup's LDS copy of it).
__device__ inline double gh_normal_from(const double *__restrict__ wt, const double *__restrict__ f,
                                        const uint64_t *__restrict__ kt, uint64_t key, uint64_t d) {
  const double inv53 = 1.1102230246251565e-16;  // 2^-53
  double x = 0.0;
  for (uint64_t a = 0; a < 64; ++a) {
    const uint64_t c = d + (a << 40);
    const uint64_t w0 = gh_word(key, 0, c);
    const int i = (int)(w0 & 0xFF);
    const bool neg = (w0 >> 8) & 1;
    const uint64_t u = w0 >> 11;
    x = (double)u * wt[i];
    if (u < kt[i]) return neg ? -x : x;
    if (i == 0) {
      const double u1 = (double)((gh_word(key, GH_LANE1, c) >> 11) + 1) * inv53;
      const double u2 = (double)((gh_word(key, GH_LANE2, c) >> 11) + 1) * inv53;
      const double xt = -det_log(u1) * (1.0 / GH_R);
      const double yt = -det_log(u2);
      if (yt + yt > xt * xt) return neg ? -(GH_R + xt) : (GH_R + xt);
    } else {
      const double u1 = (double)(gh_word(key, GH_LANE1, c) >> 11) * inv53;
      const double y = f[i] + u1 * (f[i + 1] - f[i]);
      if (y < det_exp(-0.5 * (x * x))) return neg ? -x : x;
    }
  }
  return x;
}

__device__ inline double gh_normal(uint64_t key, uint64_t d) {
  return gh_normal_from(g_zig.wt, g_zig.f, g_zig.kt, key, d);
}

// The first attempt's fast accept alone: true and the normal in v when the
// draw's first word lands inside its layer's rectangle, as gh_normal_from's
// first iteration decides; false otherwise, when the full routine is needed.
// The mix's argument z = key + GAMMA (d + 1) comes in already formed, and a
// layer's edge and bound are read together.
struct ZigPair {
  double wt;
  uint64_t kt;
};

__device__ inline bool gh_normal_fast_z(const ZigPair *__restrict__ tab, uint64_t z, double &v) {
  const uint64_t w0 = gh_mix(z);
  const int i = (int)(w0 & 0xFF);
  const bool neg = (w0 >> 8) & 1;
  const uint64_t u = w0 >> 11;
  const ZigPair p = tab[i];
  const double x = (double)u * p.wt;
  v = neg ? -x : x;
  return u < p.kt;
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void ghost_kernel(T *__restrict__ dst, int64_t ld, int64_t row0,
                                                    int64_t nrows, int64_t col0, int64_t ncols, uint64_t seed) {
  const int64_t n = nrows * ncols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int64_t i = e % nrows;
    const int64_t c = e / nrows;
    const int64_t row = row0 + i;
    const uint64_t key = gh_colkey(seed, col0 + c);
    if (CPLX) {
      const double re = gh_normal(key, 2 * (uint64_t)row) * 0.70710678118654752;
      const double im = gh_normal(key, 2 * (uint64_t)row + 1) * 0.70710678118654752;
      store_elem<T, true>(dst, i + c * ld, (T)re, (T)im);
    } else {
      store_elem<T, false>(dst, i + c * ld, (T)gh_normal(key, (uint64_t)row), T(0));
    }
  }
}

// The fill on a 1-D grid of (column, GH_ROWS-row chunk) workgroups, for fewer
// than 2^31 rows and workgroups. Each thread owns GH_E entries, GH_ROWS / GH_E
// rows apart. It takes every draw's first fast accept and stores the entries;
// a draw that misses it (about 1 in 100: a wedge or tail test, or a retry)
// goes into a workgroup queue in LDS, and after a barrier the threads resolve
// the queued draws one each with the full routine, storing that part again.
// (Resolved by their own threads, a wave would loop over the union of its
// lanes' misses with one or two lanes active.) The values are gh_normal's:
// the fast accept is its first test, and a missed draw is recomputed by it
// from the first attempt. The column key is computed once per thread; the
// mix's argument key + GAMMA (d + 1) is formed once and then stepped by
// constant multiples of GAMMA (the same sums mod 2^64 as the products); the
// ziggurat tables are read from an LDS copy, a layer's edge and bound as one
// pair.
constexpr int GH_E = 8;
constexpr int GH_ROWS = 256 * GH_E;

template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void ghost_rows_kernel(T *__restrict__ dst, int64_t ld, int64_t row0,
                                                         uint32_t nrows, DivU32 chunks, int64_t col0,
                                                         uint64_t seed) {
  constexpr int NP = CPLX ? 2 : 1;
  __shared__ double s_wt[256];
  __shared__ double s_f[257];
  __shared__ uint64_t s_kt[256];
  __shared__ ZigPair s_tab[256];
  // Queue entries: (draw bit) * 256 + thread; every draw of the workgroup
  // fits, so no miss can overflow it.
  __shared__ uint32_t s_queue[256 * GH_E * NP];
  __shared__ uint32_t s_count;
  const int t = threadIdx.x;
  s_wt[t] = g_zig.wt[t];
  s_kt[t] = g_zig.kt[t];
  s_tab[t].wt = g_zig.wt[t];
  s_tab[t].kt = g_zig.kt[t];
  s_f[t] = g_zig.f[t];
  if (t == 0) {
    s_f[256] = g_zig.f[256];
    s_count = 0;
  }
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t c = divu32(b, chunks);
  const uint32_t base = (b - c * chunks.d) * (uint32_t)GH_ROWS;
  const uint64_t key = gh_colkey(seed, col0 + (int64_t)c);
  T *out = dst + NP * (int64_t)c * ld;
  const double scale = CPLX ? 0.70710678118654752 : 1.0;
  // Draw d of row r is d = 2 r + part (complex) or r (real); its mix
  // argument is key + GAMMA (d + 1). Entry j of this thread is row
  // row0 + base + 256 j + t, so entry j + 1's argument is entry j's plus
  // GAMMA * 256 NP.
  const uint64_t first_row = (uint64_t)(row0 + (int64_t)(base + (uint32_t)t));
  uint64_t z = key + GH_GAMMA * ((CPLX ? 2 * first_row : first_row) + 1);
  const uint64_t zstep = GH_GAMMA * (uint64_t)(NP * 256);
  uint32_t pending = 0;
#pragma unroll
  for (int j = 0; j < GH_E; ++j) {
    const uint32_t i = base + (uint32_t)(j * 256 + t);
    const uint64_t zj = z;
    z += zstep;
    if (i < nrows) {
      if (CPLX) {
        double re, im;
        const bool okr = gh_normal_fast_z(s_tab, zj, re);
        const bool oki = gh_normal_fast_z(s_tab, zj + GH_GAMMA, im);
        pending |= (okr ? 0u : 1u) << (2 * j);
        pending |= (oki ? 0u : 1u) << (2 * j + 1);
        store_elem<T, true>(out, i, (T)(re * scale), (T)(im * scale));
      } else {
        double v;
        const bool ok = gh_normal_fast_z(s_tab, zj, v);
        pending |= (ok ? 0u : 1u) << j;
        store_elem<T, false>(out, i, (T)v, T(0));
      }
    }
  }
  if (pending) {
    uint32_t slot = atomicAdd(&s_count, (uint32_t)__builtin_popcount(pending));
    while (pending) {
      const int bit = __builtin_ctz(pending);
      pending &= pending - 1;
      s_queue[slot++] = (uint32_t)bit * 256u + (uint32_t)t;
    }
  }
  // The barrier also orders each first store before the store that replaces
  // one part of it.
  __syncthreads();
  const uint32_t nq = s_count;
  for (uint32_t q = (uint32_t)t; q < nq; q += 256u) {
    const uint32_t code = s_queue[q];
    const int bit = (int)(code >> 8);
    const int j = CPLX ? bit >> 1 : bit;
    const int part = CPLX ? bit & 1 : 0;
    const uint32_t i = base + (uint32_t)(j * 256) + (code & 255u);
    const uint64_t row = (uint64_t)(row0 + (int64_t)i);
    const uint64_t d = CPLX ? 2 * row + (uint64_t)part : row;
    out[NP * (int64_t)i + part] = (T)(gh_normal_from(s_wt, s_f, s_kt, key, d) * scale);
  }
}

static std::mutex g_zig_lock;
static bool g_zig_ready[64];

static int ensure_tables() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(FHIP_E_ARGUMENT, "fhip_ghost_randn: device index %d out of range", dev);
  std::lock_guard<std::mutex> guard(g_zig_lock);
  if (g_zig_ready[dev]) return FHIP_OK;
  static ZigTables host;
  static bool built = false;
  if (!built) {
    build_tables(host);
    built = true;
  }
  e = hipMemcpyToSymbol(HIP_SYMBOL(g_zig), &host, sizeof(ZigTables), 0, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyToSymbol(ziggurat tables)");
  g_zig_ready[dev] = true;
  return FHIP_OK;
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_ghost_randn(void *dst, int type, int64_t ld, int64_t row0, int64_t nrows, int64_t col0,
                                int64_t ncols, uint64_t seed, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_ghost_randn: unknown eltype");
  if (row0 < 0 || nrows < 0 || ncols < 0) return fail(FHIP_E_ARGUMENT, "fhip_ghost_randn: negative dimension");
  if (nrows == 0 || ncols == 0) return FHIP_OK;
  if (col0 < 1) return fail(FHIP_E_ARGUMENT, "fhip_ghost_randn: column indices are 1-based, got %lld", (long long)col0);
  if (!dst) return fail(FHIP_E_ARGUMENT, "fhip_ghost_randn: null pointer");
  if (ld < nrows) return fail(FHIP_E_ARGUMENT, "fhip_ghost_randn: leading dimension smaller than the rows");
  int st = ensure_tables();
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int64_t chunks = cdiv64(nrows, GH_ROWS);
  if (nrows < (int64_t(1) << 31) && chunks * ncols < (int64_t(1) << 31)) {
    const dim3 grid((unsigned)(chunks * ncols));
    const DivU32 dc = make_divu32((uint32_t)chunks);
    switch (type) {
      case FHIP_C64:
        hipLaunchKernelGGL((ghost_rows_kernel<double, true>), grid, dim3(256), 0, s, (double *)dst, ld, row0,
                           (uint32_t)nrows, dc, col0, seed);
        break;
      case FHIP_C32:
        hipLaunchKernelGGL((ghost_rows_kernel<float, true>), grid, dim3(256), 0, s, (float *)dst, ld, row0,
                           (uint32_t)nrows, dc, col0, seed);
        break;
      case FHIP_F64:
        hipLaunchKernelGGL((ghost_rows_kernel<double, false>), grid, dim3(256), 0, s, (double *)dst, ld, row0,
                           (uint32_t)nrows, dc, col0, seed);
        break;
      case FHIP_F32:
        hipLaunchKernelGGL((ghost_rows_kernel<float, false>), grid, dim3(256), 0, s, (float *)dst, ld, row0,
                           (uint32_t)nrows, dc, col0, seed);
        break;
    }
    return check_launch("ghost_rows_kernel");
  }
  int64_t g = cdiv64(nrows * ncols, 256);
  if (g > 262144) g = 262144;
  switch (type) {
    case FHIP_C64:
      hipLaunchKernelGGL((ghost_kernel<double, true>), dim3((unsigned)g), dim3(256), 0, s, (double *)dst, ld, row0,
                         nrows, col0, ncols, seed);
      break;
    case FHIP_C32:
      hipLaunchKernelGGL((ghost_kernel<float, true>), dim3((unsigned)g), dim3(256), 0, s, (float *)dst, ld, row0,
                         nrows, col0, ncols, seed);
      break;
    case FHIP_F64:
      hipLaunchKernelGGL((ghost_kernel<double, false>), dim3((unsigned)g), dim3(256), 0, s, (double *)dst, ld, row0,
                         nrows, col0, ncols, seed);
      break;
    case FHIP_F32:
      hipLaunchKernelGGL((ghost_kernel<float, false>), dim3((unsigned)g), dim3(256), 0, s, (float *)dst, ld, row0,
                         nrows, col0, ncols, seed);
      break;
  }
  return check_launch("ghost_kernel");
}

/* Checksum of the ziggurat tables as this library builds them on the host,
 * so a host-only test can compare them with an independent construction. */
extern "C" int fhip_ghost_tables_checksum(uint64_t *out) {
  if (!out) return fail(FHIP_E_ARGUMENT, "fhip_ghost_tables_checksum: null output pointer");
  ZigTables z;
  build_tables(z);
  uint64_t h = 0;
  for (int i = 0; i < 256; ++i) {
    h = gh_mix(h ^ gh_double_to_bits(z.wt[i]));
    h = gh_mix(h ^ gh_double_to_bits(z.f[i]));
    h = gh_mix(h ^ z.kt[i]);
  }
  *out = h;
  return FHIP_OK;
}
