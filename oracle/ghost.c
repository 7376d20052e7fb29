/*
 * The ghost-panel stream on the CPU: the oracle for the device fill.
 *
 * TEST INFRASTRUCTURE (see oracle/__init__.py). This restates the stream the
 * backend specifies for GhostPanels (the reference's contract,
 * src/ghosts.jl:53-63 and :97-103: column col depends on (seed, col) alone,
 * standard normal entries, complex parts of variance 1/2), written
 * independently of the device code:
 *
 *   mix(z)        SplitMix64 finalizer (Steele, Lea, Flood 2014).
 *   colkey        mix(mix(seed ^ 0x6A09E667F3BCC909) + GAMMA * col), col 1-based.
 *   word(k,L,c)   mix((k ^ L) + GAMMA * (c + 1)), lanes L in {0, LANE1, LANE2}.
 *   draw d        complex entry (row, part): d = 2 row + part; real: d = row.
 *   attempt a     counter c = d + a * 2^40.
 *   normal        256-layer ziggurat (Marsaglia and Tsang 2000), r and v as
 *                 published; layer from the low 8 bits, sign from bit 8,
 *                 53-bit uniform from the top bits; tail by Marsaglia's
 *                 exponential method, wedge by the density test.
 *   exp, log      fixed sequences of IEEE adds, multiplies and fmas
 *                 (Cody-Waite + Taylor; atanh series with a Newton
 *                 reciprocal), compiled with contraction off.
 *
 * Bit parity with the Julia stream the reference uses (hash + Xoshiro256++ +
 * randn!) is unpinned: no Julia here.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define GAMMA 0x9E3779B97F4A7C15ULL
#define SEEDK 0x6A09E667F3BCC909ULL
#define LANE1 0xB7E151628AED2A6BULL
#define LANE2 0x243F6A8885A308D3ULL
static const double ZIG_R = 3.6541528853610088;
static const double ZIG_V = 4.92867323399e-3;
static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;

uint64_t oracle_mix(uint64_t z) {
  z ^= z >> 30;
  z *= 0xBF58476D1CE4E5B9ULL;
  z ^= z >> 27;
  z *= 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return z;
}

static double as_double(uint64_t u) {
  double d;
  memcpy(&d, &u, 8);
  return d;
}

static uint64_t as_bits(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}

double oracle_exp(double x) {
  static const double inv_fact[] = {1.0,         1.0,          1.0 / 2,        1.0 / 6,        1.0 / 24,
                                    1.0 / 120,   1.0 / 720,    1.0 / 5040,     1.0 / 40320,    1.0 / 362880,
                                    1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800};
  if (x < -745.0) return 0.0;
  double t = x * 1.4426950408889634;
  double n = (t + 6755399441055744.0) - 6755399441055744.0;
  double r = fma(-n, LN2_HI, x);
  r = fma(-n, LN2_LO, r);
  double p = inv_fact[13];
  for (int j = 12; j >= 0; --j) p = fma(p, r, inv_fact[j]);
  int64_t k = (int64_t)n;
  if (k < -1000) {
    p = p * as_double((uint64_t)(1023 - 600) << 52);
    k += 600;
  }
  return p * as_double((uint64_t)(k + 1023) << 52);
}

double oracle_log(double x) {
  uint64_t b = as_bits(x);
  int64_t e = (int64_t)((b >> 52) & 0x7FF) - 1023;
  double m = as_double((b & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double d = 2.0 + f;
  double y = 0.5;
  for (int i = 0; i < 6; ++i) {
    double err = fma(-d, y, 1.0);
    y = fma(y, err, y);
  }
  double s = f * y;
  double z = s * s;
  double q = 1.0 / 23.0;
  for (int j = 21; j >= 1; j -= 2) q = fma(q, z, 1.0 / (double)j);
  double lnm = 2.0 * s * q;
  double ed = (double)e;
  return fma(ed, LN2_HI, fma(ed, LN2_LO, lnm));
}

static double wt[256], fy[257];
static uint64_t kt[256];
static int tables_ready = 0;

static double density(double v) { return oracle_exp(-0.5 * (v * v)); }

static void build(void) {
  double x[257];
  x[1] = ZIG_R;
  x[0] = ZIG_V / density(ZIG_R);
  for (int i = 1; i < 255; ++i) {
    double a = ZIG_V / x[i] + density(x[i]);
    x[i + 1] = (a >= 1.0) ? 0.0 : sqrt(-2.0 * oracle_log(a));
  }
  x[256] = 0.0;
  for (int i = 0; i < 256; ++i) {
    wt[i] = x[i] / 9007199254740992.0;
    fy[i] = density(x[i]);
    double ratio = (i == 0 ? ZIG_R : x[i + 1]) / x[i];
    kt[i] = (uint64_t)(ratio * 9007199254740992.0);
  }
  fy[256] = 1.0;
  tables_ready = 1;
}

/* Same checksum the library reports for its host-built tables. */
uint64_t oracle_tables_checksum(void) {
  if (!tables_ready) build();
  uint64_t h = 0;
  for (int i = 0; i < 256; ++i) {
    h = oracle_mix(h ^ as_bits(wt[i]));
    h = oracle_mix(h ^ as_bits(fy[i]));
    h = oracle_mix(h ^ kt[i]);
  }
  return h;
}

/* Layer edges, for the area identity check: out[0..256]. */
void oracle_zig_edges(double *out) {
  if (!tables_ready) build();
  for (int i = 0; i < 256; ++i) out[i] = wt[i] * 9007199254740992.0;
  out[256] = 0.0;
}

static uint64_t word(uint64_t key, uint64_t lane, uint64_t c) { return oracle_mix((key ^ lane) + GAMMA * (c + 1)); }

uint64_t oracle_colkey(uint64_t seed, int64_t col) { return oracle_mix(oracle_mix(seed ^ SEEDK) + GAMMA * (uint64_t)col); }

double oracle_normal(uint64_t key, uint64_t d) {
  if (!tables_ready) build();
  const double inv53 = 1.0 / 9007199254740992.0;
  double x = 0.0;
  for (uint64_t a = 0; a < 64; ++a) {
    uint64_t c = d + (a << 40);
    uint64_t w0 = word(key, 0, c);
    int i = (int)(w0 & 0xFF);
    int neg = (int)((w0 >> 8) & 1);
    uint64_t u = w0 >> 11;
    x = (double)u * wt[i];
    if (u < kt[i]) return neg ? -x : x;
    if (i == 0) {
      double u1 = (double)((word(key, LANE1, c) >> 11) + 1) * inv53;
      double u2 = (double)((word(key, LANE2, c) >> 11) + 1) * inv53;
      double xt = -oracle_log(u1) * (1.0 / ZIG_R);
      double yt = -oracle_log(u2);
      if (yt + yt > xt * xt) return neg ? -(ZIG_R + xt) : (ZIG_R + xt);
    } else {
      double u1 = (double)(word(key, LANE1, c) >> 11) * inv53;
      double y = fy[i] + u1 * (fy[i + 1] - fy[i]);
      if (y < density(x)) return neg ? -x : x;
    }
  }
  return x;
}

/* Rows row0 .. row0+nrows-1 of columns col0 .. col0+ncols-1 (1-based) of a
 * complex ghost matrix into out (column major, ld = nrows, interleaved). */
void oracle_ghost_complex(double *out, int64_t row0, int64_t nrows, int64_t col0, int64_t ncols, uint64_t seed) {
  for (int64_t c = 0; c < ncols; ++c) {
    uint64_t key = oracle_colkey(seed, col0 + c);
    for (int64_t i = 0; i < nrows; ++i) {
      uint64_t row = (uint64_t)(row0 + i);
      double re = oracle_normal(key, 2 * row) * 0.70710678118654752;
      double im = oracle_normal(key, 2 * row + 1) * 0.70710678118654752;
      out[2 * (i + c * nrows)] = re;
      out[2 * (i + c * nrows) + 1] = im;
    }
  }
}

void oracle_ghost_real(double *out, int64_t row0, int64_t nrows, int64_t col0, int64_t ncols, uint64_t seed) {
  for (int64_t c = 0; c < ncols; ++c) {
    uint64_t key = oracle_colkey(seed, col0 + c);
    for (int64_t i = 0; i < nrows; ++i) out[i + c * nrows] = oracle_normal(key, (uint64_t)(row0 + i));
  }
}
