// Shared internals of libfunicular_hip.so: error helpers, element traits and
// the gfx950 matrix-core wrappers the kernels are built from.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/funicular_hip.h"

namespace fhip {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t err, const char *what);
int check_launch(const char *what);
size_t eltype_size(int type);

inline bool is_complex(int type) { return type == FHIP_C32 || type == FHIP_C64; }
inline bool is_double(int type) { return type == FHIP_F64 || type == FHIP_C64; }
inline bool valid_type(int type) { return type >= FHIP_F32 && type <= FHIP_C64; }

inline int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Unsigned 32-bit division by a divisor fixed for a launch, as one
// multiply-high, a subtract and two shifts (Granlund and Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, Fig. 4.1).
// With l = ceil(log2 d): m = floor(2^32 (2^l - d) / d) + 1 (< 2^32), and for
// every 32-bit n, n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) with
// t = mulhi(m, n).
struct DivU32 {
  uint32_t d, m, s1, s2;
};

inline DivU32 make_divu32(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (uint64_t(1) << l) < d) ++l;
  DivU32 r;
  r.d = d;
  r.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  r.s1 = l < 1 ? l : 1;
  r.s2 = l > 1 ? l - 1 : 0;
  return r;
}

__device__ inline uint32_t divu32(uint32_t n, const DivU32 &f) {
  const uint32_t t = __umulhi(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Hardware dispatches workgroup b of a 1-D grid to XCD b % 8. This renumbers
// b so that each XCD takes one contiguous range of the nb logical ids (XCD x
// takes nb / 8 of them, one more when x < nb % 8), so neighbouring work shares
// that XCD's L2.
__device__ inline uint32_t xcd_contiguous(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, local = b >> 3, per = nb >> 3, extra = nb & 7u;
  return x * per + (x < extra ? x : extra) + local;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core product per call, D = A*B + C, on a 64-wide wave.
// Operand layout: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// element r of the result holds D[row(l / 16, r)][l % 16]. The row map differs
// between the two forms (measured on gfx950): the f64 form interleaves the
// four lane groups, row = l/16 + 4 r; the f32 form gives each lane group four
// consecutive rows, row = 4 (l/16) + r.
template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef v4d acc;
  __device__ static inline acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return lk + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef v4f acc;
  __device__ static inline acc zero() { return acc{0.0f, 0.0f, 0.0f, 0.0f}; }
  __device__ static inline acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return 4 * lk + r; }
};

// How the complex product of one K step combines its four real products.
enum ProductMode {
  PROD_CONJ_A = 0,  // acc += conj(a) * b   (Gram: A^H B)
  PROD_ADD = 1,     // acc += a * b         (D = A C)
  PROD_SUB = 2      // acc -= a * b         (trailing update of a solve)
};

// K steps s = s_begin, s_begin + s_stride, ... < s_end of a wave's MB x NB
// block tile. As/Bs are LDS planes laid out [k][m] and [k][n] with row strides
// lda_s/ldb_s; plane 1 holds imaginary parts when CPLX. The wave's tile starts
// at column mi0 of As and column nj0 of Bs.
template <typename T, bool CPLX, int MB, int NB, int MODE>
__device__ inline void mma_steps(const T *__restrict__ As_re, const T *__restrict__ As_im, int lda_s,
                                 int mi0, const T *__restrict__ Bs_re, const T *__restrict__ Bs_im,
                                 int ldb_s, int nj0, int s_begin, int s_end, int s_stride,
                                 typename Mfma<T>::acc (&acc_re)[MB][NB],
                                 typename Mfma<T>::acc (&acc_im)[MB][NB]) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  for (int s = s_begin; s < s_end; s += s_stride) {
    const int kr = 4 * s + lk;
    T ar[MB], ai[MB], br[NB], bi[NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      ar[mb] = As_re[kr * lda_s + mi0 + 16 * mb + li];
      if (CPLX) ai[mb] = As_im[kr * lda_s + mi0 + 16 * mb + li];
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      br[nb] = Bs_re[kr * ldb_s + nj0 + 16 * nb + li];
      if (CPLX) bi[nb] = Bs_im[kr * ldb_s + nj0 + 16 * nb + li];
    }
    if (!CPLX) {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const T a = (MODE == PROD_SUB) ? -ar[mb] : ar[mb];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc_re[mb][nb] = Mfma<T>::mma(a, br[nb], acc_re[mb][nb]);
      }
    } else {
      // First half: the products with the real part of a.
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const T a = (MODE == PROD_SUB) ? -ar[mb] : ar[mb];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          acc_re[mb][nb] = Mfma<T>::mma(a, br[nb], acc_re[mb][nb]);
          acc_im[mb][nb] = Mfma<T>::mma(a, bi[nb], acc_im[mb][nb]);
        }
      }
      // Second half: the products with the imaginary part of a.
      //   conj(a) b: re += ai bi, im -= ai br
      //   a b:       re -= ai bi, im += ai br
      //   -a b:      re += ai bi, im -= ai br
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const T p = (MODE == PROD_ADD) ? -ai[mb] : ai[mb];
        const T q = (MODE == PROD_ADD) ? ai[mb] : -ai[mb];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          acc_re[mb][nb] = Mfma<T>::mma(p, bi[nb], acc_re[mb][nb]);
          acc_im[mb][nb] = Mfma<T>::mma(q, br[nb], acc_im[mb][nb]);
        }
      }
    }
  }
}

// LDS column swizzle for FP64 [k][column] operand planes whose rows are 32 or
// more columns wide: column c of row kr sits at c ^ lds_swizzle(kr). Bit 4
// flips between even and odd rows, so the two rows a half wave reads for one
// matrix-core operand land in different halves of the banks; the low four
// bits take the row's own low four bits, so the 16 rows of a slice that a
// staging store writes (16 lanes, one column; or 32 lanes, two columns) fall
// on distinct bank pairs whether 16 or 32 lanes of a 64-bit access are served
// together. (An earlier form put only 8 distinct values in the low bits,
// which left the 16-lane stores 2-way conflicted.)
__device__ inline int lds_swizzle(int kr) { return (kr & 15) | ((kr & 1) << 4); }

// The 4-byte form, for FP32 planes exactly 64 columns wide: column c of row kr
// sits at c ^ lds_swizzle4(kr). A matrix-core operand read (lanes li + 16 lk
// reading row 4 s + lk, columns 16-aligned + li) then puts its four rows in
// four different 16-bank groups, and a staging store (lanes kk + 16 cc writing
// row kk, column 4-aligned + cc, 16 rows per slice) touches 64 different banks;
// both hold for 64 banks serving a whole wave and for 32 banks serving half a
// wave. Padding instead (rows of 68) leaves the reads 4-way conflicted.
__device__ inline int lds_swizzle4(int kr) { return ((kr & 3) << 4) | (kr & 2) | (((kr >> 2) & 3) << 2); }

// Layout kinds: 0 plain (padded rows), 1 lds_swizzle, 2 lds_swizzle4.
template <int SWZ>
__device__ inline int swizzle_of(int kr) {
  return SWZ == 1 ? lds_swizzle(kr) : (SWZ == 2 ? lds_swizzle4(kr) : 0);
}

// The same K steps for complex operands with three real products per complex
// product instead of four (the 3M method; Higham, "Stability of a method for
// multiplying complex matrices with three real matrix multiplications", SIMAX
// 1992). Per element pair the products are
//     p1 += ar br,   p2 += ai bi,   p3 += (ar + ai) (br -/+ bi),
// the sign being - for conj(a) b and + for a b; combine3m() turns the three
// sums into the real and imaginary parts. PROD_SUB negates a, so the sums
// count down from whatever the accumulators started at. The error bound is
// normwise, |error| <= c u |a| |b| per product, the same order as the
// four-product form's for these column-scaled sums.
template <typename T, int MB, int NB, int MODE, int SWZ = 0>
__device__ inline void mma_steps_3m(const T *__restrict__ As_re, const T *__restrict__ As_im, int lda_s,
                                    int mi0, const T *__restrict__ Bs_re, const T *__restrict__ Bs_im,
                                    int ldb_s, int nj0, int s_begin, int s_end, int s_stride,
                                    typename Mfma<T>::acc (&p1)[MB][NB], typename Mfma<T>::acc (&p2)[MB][NB],
                                    typename Mfma<T>::acc (&p3)[MB][NB], unsigned nb_mask = ~0u) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  for (int s = s_begin; s < s_end; s += s_stride) {
    const int kr = 4 * s + lk;
    const int sw = swizzle_of<SWZ>(kr);
    T ar[MB], ai[MB], as[MB], br[NB], bi[NB], bt[NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      ar[mb] = As_re[kr * lda_s + ((mi0 + 16 * mb + li) ^ sw)];
      ai[mb] = As_im[kr * lda_s + ((mi0 + 16 * mb + li) ^ sw)];
      as[mb] = ar[mb] + ai[mb];
      if (MODE == PROD_SUB) {
        ar[mb] = -ar[mb];
        ai[mb] = -ai[mb];
        as[mb] = -as[mb];
      }
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      if (!((nb_mask >> nb) & 1u)) continue;
      br[nb] = Bs_re[kr * ldb_s + ((nj0 + 16 * nb + li) ^ sw)];
      bi[nb] = Bs_im[kr * ldb_s + ((nj0 + 16 * nb + li) ^ sw)];
      bt[nb] = (MODE == PROD_CONJ_A) ? br[nb] - bi[nb] : br[nb] + bi[nb];
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        // nb_mask (wave-uniform) leaves out column tiles whose slice of B is
        // known to be zero.
        if (!((nb_mask >> nb) & 1u)) continue;
        p1[mb][nb] = Mfma<T>::mma(ar[mb], br[nb], p1[mb][nb]);
        p2[mb][nb] = Mfma<T>::mma(ai[mb], bi[nb], p2[mb][nb]);
        p3[mb][nb] = Mfma<T>::mma(as[mb], bt[nb], p3[mb][nb]);
      }
  }
}

// Real and imaginary parts from the three sums of mma_steps_3m:
//   conj(a) b:  re = p1 + p2,  im = p1 - p2 - p3;
//   a b, -a b:  re = p1 - p2,  im = p3 - p1 - p2.
template <int MODE, typename V>
__device__ inline void combine3m(const V &p1, const V &p2, const V &p3, V &re, V &im) {
  if (MODE == PROD_CONJ_A) {
    re = p1 + p2;
    im = (p1 - p2) - p3;
  } else {
    re = p1 - p2;
    im = p3 - (p1 + p2);
  }
}

template <typename T>
struct Pair;
template <>
struct Pair<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};

// Raw buffer loads. A descriptor covers nbytes from base; a load whose byte
// offset lies past nbytes returns zero instead of touching memory, so a
// descriptor ending where a slice's valid columns end zeroes the columns
// past k without a per-lane test. The per-lane offset is a 32-bit VGPR and
// the descriptor four SGPRs, so a slice's loads need no 64-bit address
// arithmetic. Offsets are checked on the VGPR part alone: the SGPR offset
// argument stays 0 and a slice moves its descriptor's base instead.
// (0x00020000 is the gfx9 word 3 for raw, untyped access.)
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)nbytes, 0x00020000);
}

template <typename P>
__device__ inline P buffer_load_pair(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ inline Pair<double>::type buffer_load_pair<Pair<double>::type>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(Pair<double>::type, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
template <>
__device__ inline Pair<float>::type buffer_load_pair<Pair<float>::type>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(Pair<float>::type, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// Loads one element of a real or complex matrix as (re, im) in type T. A
// complex element is read as one aligned pair (ComplexF64 and ComplexF32 are
// aligned to their full size).
template <typename TIn, typename T, bool CPLX>
__device__ inline void load_elem(const TIn *__restrict__ p, int64_t idx, bool ok, T &re, T &im) {
  if (!ok) {
    re = T(0);
    im = T(0);
    return;
  }
  if (CPLX) {
    typedef typename Pair<TIn>::type P2;
    const P2 v = reinterpret_cast<const P2 *>(p)[idx];
    re = T(v.x);
    im = T(v.y);
  } else {
    re = T(p[idx]);
    im = T(0);
  }
}

// Branch-free load of one element as it is stored, for staging registers: a
// lane outside the matrix reads element 0 instead (always a valid address),
// and from_raw() zeroes it. Keeping the conversion to the compute type out of
// the load lets the load stay in flight until the value is staged; a
// conversion next to a guarded load makes the compiler wait for each load on
// the spot.
template <typename TIn, bool CPLX>
__device__ inline typename Pair<TIn>::type load_raw(const TIn *__restrict__ p, int64_t idx, bool ok) {
  typedef typename Pair<TIn>::type P2;
  const int64_t i = ok ? idx : 0;
  if (CPLX) return reinterpret_cast<const P2 *>(p)[i];
  P2 v;
  v.x = p[i];
  v.y = TIn(0);
  return v;
}

template <typename T, typename P2>
__device__ inline void from_raw(const P2 &v, bool ok, T &re, T &im) {
  re = ok ? T(v.x) : T(0);
  im = ok ? T(v.y) : T(0);
}

// Stores one element of a real or complex matrix.
template <typename TOut, bool CPLX>
__device__ inline void store_elem(TOut *__restrict__ p, int64_t idx, TOut re, TOut im) {
  if (CPLX) {
    typedef typename Pair<TOut>::type P2;
    P2 v;
    v.x = re;
    v.y = im;
    reinterpret_cast<P2 *>(p)[idx] = v;
  } else {
    p[idx] = re;
  }
}

}  // namespace fhip
