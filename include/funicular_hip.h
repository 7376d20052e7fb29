/*
 * funicular_hip.h -- C ABI of libfunicular_hip.so, the MI355X (gfx950) backend
 * behind Funicular.jl's device contract.
 *
 * Every entry point returns an int status: FHIP_OK (0) on success, otherwise one
 * of the FHIP_E_* codes below. The message that goes with the last failure on
 * the calling thread is fhip_last_error(). Nothing here aborts the process and
 * nothing synchronizes unless its name says so (the *_sync calls and
 * fhip_device_sync). Streams are passed explicitly everywhere; a NULL stream is
 * the null stream.
 *
 * Matrices are column major. A matrix argument is a base pointer plus a leading
 * dimension `ld` counted in elements, so a row range or a column range of a
 * larger buffer is passed by offsetting the pointer and keeping `ld`. Complex
 * elements are interleaved (re, im) pairs, which is Julia's ComplexF64 and
 * ComplexF32 layout. Element types are FHIP_F32, FHIP_F64, FHIP_C32 and
 * FHIP_C64. Complex scalars travel as two doubles.
 *
 * Each declaration names the reference interface it replaces, as
 * file:line in PaulVirally/Funicular.jl v0.3.0.
 */
#ifndef FUNICULAR_HIP_H
#define FUNICULAR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FHIP_OK 0
#define FHIP_E_ARGUMENT 1          /* Julia ArgumentError */
#define FHIP_E_DIMENSION 2         /* Julia DimensionMismatch */
#define FHIP_E_HIP 3               /* the HIP runtime reported an error */
#define FHIP_E_WORKSPACE 4         /* workspace smaller than the *_workspace_bytes query */
#define FHIP_E_UNSUPPORTED 5       /* eltype combination this library has no kernel for */

/* element types */
#define FHIP_F32 0
#define FHIP_F64 1
#define FHIP_C32 2
#define FHIP_C64 3

/* copy directions for fhip_memcpy2d_async */
#define FHIP_H2D 1
#define FHIP_D2H 2
#define FHIP_D2D 3
#define FHIP_H2H 4

/* ---------------------------------------------------------------- errors -- */

/* Message of the last failure on the calling thread ("" when none). Replaces
 * the exceptions the Julia backend throws (src/device_api.jl:161,
 * ext/FunicularCUDAExt.jl:123). */
const char *fhip_last_error(void);

/* Library version string. */
const char *fhip_version(void);

/* --------------------------------------------------------------- devices -- */

int fhip_device_count(int *count);
int fhip_set_device(int device);
/* Name of the current device (gcnArchName) into buf. */
int fhip_device_arch(char *buf, size_t len);

/* Blocks until every stream of the current device has drained.
 * Replaces sync_all (src/device_api.jl:109, ext/FunicularCUDAExt.jl:114-117). */
int fhip_device_sync(void);

/* ---------------------------------------------------------------- memory -- */

/* Device allocation, synchronized before return so that any stream may use it.
 * Replaces alloc_device (src/device_api.jl:19, ext/FunicularCUDAExt.jl:32-40). */
int fhip_malloc(void **ptr, size_t nbytes);
int fhip_free(void *ptr);

/* Page-locked host allocation. Replaces alloc_host_slab
 * (src/device_api.jl:28, ext/FunicularCUDAExt.jl:45-49); the 64-byte slab
 * alignment of src/device_api.jl:140-154 is met by the allocator itself. */
int fhip_host_alloc(void **ptr, size_t nbytes);
int fhip_host_free(void *ptr);
/* *pinned = 1 when ptr lies in page-locked host memory the runtime knows
 * (an fhip_host_alloc block), 0 for pageable memory. The check the
 * reference's CUDA suite makes with CUDA.is_pinned on a host slab
 * (test/test_cuda.jl:24-36). */
int fhip_host_is_pinned(int *pinned, const void *ptr);

/* ---------------------------------------------------------------- queues -- */

/* A non-blocking stream at the default priority. Replaces make_queue and
 * make_sweep_queues (src/device_api.jl:53,63, ext/FunicularCUDAExt.jl:71-81). */
int fhip_stream_create(void **stream);
int fhip_stream_destroy(void *stream);
/* Replaces sync_queue (src/device_api.jl:95, ext/FunicularCUDAExt.jl:104-107). */
int fhip_stream_sync(void *stream);

/* Events. disable_timing=1 matches CuEvent(EVENT_DISABLE_TIMING)
 * (ext/FunicularCUDAExt.jl:91-95); timing events are for measurement only. */
int fhip_event_create(void **event, int disable_timing);
int fhip_event_destroy(void *event);
/* Replaces record_event (src/device_api.jl:81, ext/FunicularCUDAExt.jl:91-95). */
int fhip_event_record(void *event, void *stream);
/* Device-side wait that does not block the caller. Replaces wait_event
 * (src/device_api.jl:88, ext/FunicularCUDAExt.jl:99-102). */
int fhip_stream_wait_event(void *stream, void *event);
/* Replaces sync_event (src/device_api.jl:102, ext/FunicularCUDAExt.jl:109-112). */
int fhip_event_sync(void *event);
/* *done = 1 when the event has completed, 0 otherwise. */
int fhip_event_query(void *event, int *done);
/* Milliseconds between two completed timing events. */
int fhip_event_elapsed(float *ms, void *start, void *stop);

/* ---------------------------------------------------------------- copies -- */

/* 2-D copy of `rows` lines of `width_bytes` bytes, lines `spitch`/`dpitch`
 * bytes apart, asynchronous on `stream` (synchronous when stream is NULL and
 * sync is 1). In Funicular's column-major terms a "line" is one column and
 * the pitch is the leading dimension in bytes. Replaces h2d!/d2h! and the
 * copy_region! they reach (src/device_api.jl:37,45,
 * ext/FunicularCUDAExt.jl:119-152). */
int fhip_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                        size_t width_bytes, size_t rows, int kind, void *stream);

/* ---------------------------------------------------------- diagnostics -- */

typedef struct fhip_counters {
  uint64_t device_allocs;      /* fhip_malloc calls that succeeded */
  uint64_t device_bytes;       /* bytes currently allocated by fhip_malloc */
  uint64_t host_allocs;        /* fhip_host_alloc calls that succeeded */
  uint64_t host_bytes;         /* bytes currently allocated by fhip_host_alloc */
  uint64_t h2d_bytes;          /* bytes issued host->device */
  uint64_t d2h_bytes;          /* bytes issued device->host */
  uint64_t d2d_bytes;          /* bytes issued device->device */
  uint64_t copies;             /* copy calls issued */
  uint64_t kernel_launches;    /* kernels launched by this library */
  uint64_t host_syncs;         /* blocking stream/event/device waits */
} fhip_counters;

/* Snapshot of the process-wide counters. Replaces the counters the reference
 * exposes (src/plan.jl:185, src/plan.jl:312-314, src/io.jl:141-142) and the
 * allocation discipline checks of test/test_cuda.jl:78-111. */
int fhip_get_counters(fhip_counters *out);
int fhip_reset_counters(void);

/* -------------------------------------------------------------- kernels -- */

/* Workspace bytes fhip_gram needs for an m-row product of ka by kb columns in
 * eltype `type` (compute eltype `compute`; equal to `type` except for the mixed
 * pass). `hermitian` = 1 when A and B are the same matrix. */
int fhip_gram_workspace_bytes(size_t *nbytes, int64_t m, int64_t ka, int64_t kb,
                              int type, int compute, int hermitian);

/* C <- alpha * A^H * B + beta * C with A m x ka (lda), B m x kb (ldb) and C
 * ka x kb (ldc). A zero beta overwrites C whatever it held, NaN included.
 * When A and B are the same pointer with the same ld and width the result is
 * formed from its upper triangle and is exactly Hermitian for real alpha and
 * beta. The row sum is split into chunks fixed by (m, ka, kb) alone and added
 * in a fixed order, so the bits do not depend on the stream, the workspace or
 * how many staging buffers a sweep uses.
 * `type` is the storage eltype of A, B and C; `compute` the eltype of the
 * matrix-core arithmetic: FHIP_C32 with type FHIP_C64 is the mixed pass
 * (inputs narrowed while staged, FP32 products per chunk, chunks summed in
 * FP64 into an FP64 C). Otherwise compute must equal type.
 * Replaces gram_accumulate! (src/kernels.jl:40; semantics src/kernels.jl:22-31).
 */
int fhip_gram(void *C, int64_t ldc, const void *A, int64_t lda, const void *B,
              int64_t ldb, int64_t m, int64_t ka, int64_t kb, double alpha_re,
              double alpha_im, double beta_re, double beta_im, int type, int compute,
              void *work, size_t work_bytes, void *stream);

/* Workspace bytes fhip_trsm_runn needs for a k x k factor solved in eltype
 * `compute`. */
int fhip_trsm_workspace_bytes(size_t *nbytes, int64_t k, int compute);

/* Y <- Y * R^-1 in place, R upper triangular k x k with a non-unit diagonal
 * (right side, upper, no transpose). Y is m x k (ldy), R is k x k (ldr), both
 * of eltype `type`. Rows are independent. The solve is blocked by 64 columns;
 * each call first folds the inverted diagonal blocks into the factor, kept in
 * the workspace (m = 0 does only that). `compute` equals `type`,
 * or is FHIP_C32 for a FHIP_C64 Y: the mixed first pass, which narrows Y and
 * R while staging, solves on the FP32 matrix cores and writes FP64 back.
 * Replaces rdiv_upper! (src/kernels.jl:108; semantics src/kernels.jl:90-99). */
int fhip_trsm_runn(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R,
                   int64_t ldr, int type, int compute, void *work, size_t work_bytes,
                   void *stream);

/* fhip_trsm_runn without a workspace argument, and without allocating: the
 * folded factor goes to a fixed 16 MiB device array inside the library, one
 * per GPU, held from the first call on that GPU until the process ends.
 * That bounds k to 1024 for FHIP_C64 compute (kp^2 x eltype bytes <= 16 MiB,
 * kp = k rounded up to 64); a larger k fails with FHIP_E_ARGUMENT before any
 * device work. The results are those of fhip_trsm_runn, bit for bit. Every
 * call on one GPU shares that array, so calls that overlap on one GPU (two
 * streams, or two threads, without ordering between them) race. No
 * counterpart in the reference: it is for callers that cannot allocate, such
 * as code under stream capture or a serving loop that forbids hipMalloc. */
int fhip_trsm_runn_noalloc(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr,
                           int type, int compute, void *stream);

/* Y <- Y * R^-H in place (right side, R upper triangular k x k, conjugate
 * transpose): the solve of the QR adjoint, dY = (dQ + Q N) R^-H. Same
 * blocking as fhip_trsm_runn, walking the column blocks right to left. No
 * counterpart in the reference (its cholqr2! has no adjoint); the workspace
 * is fhip_trsm_workspace_bytes(k, type). */
int fhip_trsm_runc(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr,
                   int type, void *work, size_t work_bytes, void *stream);

/* The solve of fhip_trsm_runn (adjoint = 0) or fhip_trsm_runc (adjoint = 1)
 * on another m x k block, reusing the folded factor that the last such call
 * with the same k, type and compute left in `work` (on the same stream, or
 * ordered before this one). A row sweep folds the factor on its first block
 * only; the reference's rdiv_rows! solves every block with the same R
 * (src/panelblas.jl:157-169). */
int fhip_trsm_apply(void *Y, int64_t ldy, int64_t m, int64_t k, int type, int compute, int adjoint,
                    const void *work, size_t work_bytes, void *stream);

/* R <- the upper Cholesky factor of the Hermitian (symmetric) k x k matrix G,
 * whose upper triangle alone is read; R's lower triangle is set to zero. R's
 * eltype equals G's, or narrows FHIP_C64 to FHIP_C32 (FHIP_F64 to FHIP_F32)
 * for the mixed first pass. *info (a device int) becomes 0, or the 1-based
 * column of the first pivot that is not positive and finite, as LAPACK's
 * potrf reports it; the caller reads it after the stream. Replaces the plain
 * factorization of cholesky_upper (src/panelblas.jl:148-155), which the
 * reference runs on the host; its shifted retry stays on the host. */
int fhip_potrf_upper(void *R, int64_t ldr, int rtype, const void *G, int64_t ldg, int gtype,
                     int64_t k, int *info, void *stream);

/* D <- A * C, A m x s (lda), C s x r (ldc), D m x r (ldd). D must not alias A.
 * Replaces rightmul_gemm! (src/kernels.jl:73; semantics src/kernels.jl:57-64). */
int fhip_gemm_nn(void *D, int64_t ldd, const void *A, int64_t lda, const void *C,
                 int64_t ldc, int64_t m, int64_t s, int64_t r, int type, void *stream);

/* D <- alpha A C + beta D, shapes as fhip_gemm_nn (beta = 0 does not read D;
 * the imaginary parts are ignored for real types). The update step
 * Y_j <- Y_j - Q_i R_ij of the panel-by-panel triangular solve that
 * cholqr2 of a basis larger than the host budget needs; the reference has no
 * such traversal (its row sweeps need every panel host resident,
 * src/executor.jl:66-67). */
int fhip_gemm_nn_scaled(void *D, int64_t ldd, const void *A, int64_t lda, const void *C,
                        int64_t ldc, int64_t m, int64_t s, int64_t r, double alpha_re,
                        double alpha_im, double beta_re, double beta_im, int type, void *stream);

/* dst <- convert(dst_type, src), both rows x cols on the device.
 * Replaces convert_copy! (src/kernels.jl:14-20). */
int fhip_convert_copy(void *dst, int dst_type, int64_t ldd, const void *src, int src_type,
                      int64_t lds, int64_t rows, int64_t cols, void *stream);

/* Y <- a X + b Y for rows x cols matrices of one eltype (b = 0 still reads Y).
 * The elementwise step of the reference's axpy!/scale! (src/panelblas.jl:
 * 311-331) and of the residual check Y - Q R. */
int fhip_axpby(void *Y, int64_t ldy, const void *X, int64_t ldx, int64_t rows, int64_t cols,
               double a_re, double a_im, double b_re, double b_im, int type, void *stream);

/* Y <- Y alpha for a rows x cols matrix, entry by entry as rmul! computes it:
 * a real scalar (real_scalar = 1, or a real eltype) scales both parts, a
 * complex one is the schoolbook product, each part rounded once per product
 * and sum (no fused multiply-add). At most 65535 columns per call.
 * Replaces the rmul! step of scale! (src/panelblas.jl:311-316). */
int fhip_scale(void *Y, int64_t ldy, int64_t rows, int64_t cols, double a_re, double a_im,
               int real_scalar, int type, void *stream);

/* Y <- alpha Y + beta for a rows x cols matrix, entry by entry: the product
 * as fhip_scale forms it (real_scalar as there), then beta added, each part
 * rounded once per operation (no fused multiply-add). At most 65535 columns
 * per call. The broadcasts a foreachpanel closure applies to its device
 * buffer (src/executor.jl:241-267): panel .= a .* panel .+ b
 * (test/test_executor.jl:30-31); panel .= b is fhip_fill. */
int fhip_affine(void *Y, int64_t ldy, int64_t rows, int64_t cols, double a_re, double a_im,
                int real_scalar, double b_re, double b_im, int type, void *stream);

/* Bytes of workspace fhip_dot needs for a rows x cols block. */
int fhip_dot_workspace_bytes(size_t *nbytes, int64_t rows, int64_t cols);

/* out[0..1] <- (accumulate ? out : 0) + sum over a rows x cols block of
 * conj(X) .* Y (mode 0) or |X|^2 (mode 1, Y unused, imaginary part 0), summed
 * in FP64 in an order fixed by (rows, cols) alone. out is a device pointer to
 * (re, im) FP64; work holds fhip_dot_workspace_bytes. At most 65535 columns.
 * Replaces the per-panel broadcasts of dot and norm and their final sum
 * (src/panelblas.jl:333-370), without the N x w accumulator. */
int fhip_dot(double *out, const void *X, int64_t ldx, const void *Y, int64_t ldy, int64_t rows,
             int64_t cols, int type, int mode, int accumulate, void *work, size_t work_bytes,
             void *stream);

/* out[i] (+)= Re sum_j conj(A[i, j]) B[i, j] for i < rows (out is FP64,
 * overwritten unless accumulate = 1): the per-row reduction behind the
 * sensitivity of S = Q^H G Q to the operator's diagonal, dL/dd_i =
 * Re sum_a conj(Q_ia) (Q dS^H)_ia. No counterpart in the reference. */
int fhip_rowdot(double *out, const void *A, int64_t lda, const void *B, int64_t ldb, int64_t rows,
                int64_t cols, int type, int accumulate, void *stream);

/* dst <- value for every entry of a rows x cols matrix (value_im ignored for
 * real types). A kernel rather than a memset, so it is safe under capture. */
int fhip_fill(void *dst, int type, int64_t ld, int64_t rows, int64_t cols, double value_re,
              double value_im, void *stream);

/* Ghost-panel fill: dst[i, c] = entry (row0 + i, col0 + c) of a GhostPanels
 * matrix, for i < nrows and c < ncols; rows are 0-based and columns 1-based,
 * as the reference numbers them. Each column's stream depends on (seed, col)
 * only and every entry is addressed by its row, so the values do not depend
 * on the panel width, on which panel or row block asks, or on the matrix
 * height. Entries are standard normal; complex entries have real and
 * imaginary parts of variance 1/2. Replaces stage_host!(::GhostTier) and
 * generate_panel! (src/ghosts.jl:97-122) with generation straight into
 * device memory. */
int fhip_ghost_randn(void *dst, int type, int64_t ld, int64_t row0, int64_t nrows,
                     int64_t col0, int64_t ncols, uint64_t seed, void *stream);

/* Checksum of the ziggurat tables the ghost fill builds on the host (no device
 * work), so a host-only test can compare them with an independent
 * construction. Diagnostic companion of fhip_ghost_randn; it replaces nothing
 * in the reference. */
int fhip_ghost_tables_checksum(uint64_t *out);

/* Matrix-free operators of the benchmark configurations, applied to a panel:
 * Y <- op(X) for X, Y N x w (ldx, ldy), complex.
 *   tridiagonal: (Gx)_i = d x_i + o (x_{i-1} + x_{i+1}), indices mod N;
 *   adjoint=1 applies G^H (conjugated coefficients);
 *   diag != NULL adds diag[i] (real, length N) to the diagonal.
 * Replaces apply_operator! for a panel-capable operator (src/operator.jl:80). */
int fhip_stencil_tridiag(void *Y, int64_t ldy, const void *X, int64_t ldx, int64_t N,
                         int64_t w, double d_re, double d_im, double o_re, double o_im,
                         const double *diag, int adjoint, int type, void *stream);

/*   7-point stencil on an n1 x n2 x n3 periodic grid, linear index
 *   i + n1 (j + n2 l): (Gx) = d x + o (sum of the six neighbours). */
int fhip_stencil_helmholtz3d(void *Y, int64_t ldy, const void *X, int64_t ldx, int64_t n1,
                             int64_t n2, int64_t n3, int64_t w, double d_re, double d_im,
                             double o_re, double o_im, int adjoint, int type, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FUNICULAR_HIP_H */
