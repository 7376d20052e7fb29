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

__device__ inline double gh_normal(uint64_t key, uint64_t d) {
  const double inv53 = 1.1102230246251565e-16;  // 2^-53
  double x = 0.0;
  for (uint64_t a = 0; a < 64; ++a) {
    const uint64_t c = d + (a << 40);
    const uint64_t w0 = gh_word(key, 0, c);
    const int i = (int)(w0 & 0xFF);
    const bool neg = (w0 >> 8) & 1;
    const uint64_t u = w0 >> 11;
    x = (double)u * g_zig.wt[i];
    if (u < g_zig.kt[i]) return neg ? -x : x;
    if (i == 0) {
      const double u1 = (double)((gh_word(key, GH_LANE1, c) >> 11) + 1) * inv53;
      const double u2 = (double)((gh_word(key, GH_LANE2, c) >> 11) + 1) * inv53;
      const double xt = -det_log(u1) * (1.0 / GH_R);
      const double yt = -det_log(u2);
      if (yt + yt > xt * xt) return neg ? -(GH_R + xt) : (GH_R + xt);
    } else {
      const double u1 = (double)(gh_word(key, GH_LANE1, c) >> 11) * inv53;
      const double y = g_zig.f[i] + u1 * (g_zig.f[i + 1] - g_zig.f[i]);
      if (y < det_exp(-0.5 * (x * x))) return neg ? -x : x;
    }
  }
  return x;
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
  int64_t g = cdiv64(nrows * ncols, 256);
  if (g > 262144) g = 262144;
  hipStream_t s = (hipStream_t)stream;
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
