/* bh_blas_hip — device BLAS extension methods for the hip vector engine.
 *
 * libbh_blas_hip.so is an extension-method library in the reference's sense
 * (include/bohrium/bh_extmethod.hpp:57-125, found by core/bh_extmethod.cpp:33-90
 * through the engine's `libs` option): for each method it exports
 * `<name>_create` / `<name>_destroy` returning an ExtmethodImpl.  It replaces
 * the reference's CBLAS-backed extmethods/blas (templates/body.tpl:1-50,
 * body_func.tpl:1-30, methods.json) for the methods below, on the device:
 *
 *   blas_gemm    C := A B          A m x k, B k x n, C m x n   (body_func.tpl: alpha 1, beta 0)
 *   blas_gemmt   C := A^T B        A m x k, B m x n, C k x n
 *   blas_symm    C := A B          A m x m symmetric, read from its upper triangle
 *   blas_hemm    C := A B          A m x m Hermitian, upper triangle, diagonal taken as real (complex)
 *   blas_syrk    C := A A^T        upper triangle of C (n x n), A n x k
 *   blas_herk    C := A A^H        upper triangle, real diagonal (complex)
 *   blas_syr2k   C := A B^T + B A^T  upper triangle, A and B n x k
 *   blas_her2k   C := A B^H + B A^H  upper triangle, real diagonal (complex)
 *   blas_trmm    B := A B          A m x m unit upper triangular (diagonal and lower part not read)
 *   blas_trsm    B := A^-1 B       A m x m unit upper triangular
 *
 * The k-update methods write only C's upper triangle; its strictly lower part
 * keeps what C held.  Instruction operands are (C, A, B) and (C, A) for
 * syrk/herk, as the reference front end queues them
 * (bridge/npbackend/bohrium/blas.py:12-117); trmm/trsm write B and ignore C; the
 * front end applies alpha and beta before the call.  Matrices are row-major and
 * C-contiguous; float32 and float64 run on MFMA (v_mfma_f32_16x16x4f32,
 * v_mfma_f64_16x16x4f64), complex64/complex128 on a tiled FMA kernel.
 *
 * The bhblas_* functions are the same device routines as plain C entry points
 * (device pointers, leading dimensions in elements, the engine stream or any
 * HIP stream; they return without waiting, 0 on success, else
 * bhblas_last_error() says why) for callers outside a component stack and for
 * benchmarks.  symm/hemm/trmm use a stream-ordered device temporary of m x m
 * (trmm: also m x n) elements.
 */
#ifndef BH_BLAS_HIP_H
#define BH_BLAS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
namespace bohrium {
namespace extmethod {
class ExtmethodImpl;
}
}
typedef bohrium::extmethod::ExtmethodImpl bh_extmethod_impl;
extern "C" {
#else
typedef struct bh_extmethod_impl bh_extmethod_impl;
#endif

enum { BHBLAS_F32 = 0, BHBLAS_F64 = 1, BHBLAS_C64 = 2, BHBLAS_C128 = 3 };

/* C = op(A) * B with op(A) = A (trans_a = 0, A is m x k) or A^T (trans_a = 1, A is stored k x m);
 * B is k x n, C is m x n; all row-major with leading dimensions in elements.  Launches on `stream`
 * (a hipStream_t; NULL = the default stream) and returns without waiting.  0 on success. */
int bhblas_gemm(int dtype, int trans_a, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda,
                const void *B, int64_t ldb, void *C, int64_t ldc, void *stream);
/* The upper triangle (i <= j) of C = A * A^T, A m x k, C m x m (cblas_?syrk, methods.json "syrk"). */
int bhblas_syrk(int dtype, int64_t m, int64_t k, const void *A, int64_t lda, void *C, int64_t ldc, void *stream);
/* The upper triangle of C = A * A^H with a zero imaginary diagonal; complex only (cblas_?herk). */
int bhblas_herk(int dtype, int64_t n, int64_t k, const void *A, int64_t lda, void *C, int64_t ldc, void *stream);
/* The upper triangle of C = A B^T + B A^T, A and B n x k (cblas_?syr2k). */
int bhblas_syr2k(int dtype, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                 int64_t ldc, void *stream);
/* The upper triangle of C = A B^H + B A^H with a zero imaginary diagonal; complex only (cblas_?her2k). */
int bhblas_her2k(int dtype, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                 int64_t ldc, void *stream);
/* C = A B, A m x m symmetric given by its upper triangle, B and C m x n (cblas_?symm, left, upper). */
int bhblas_symm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                int64_t ldc, void *stream);
/* bhblas_symm without any allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and result as bhblas_symm; the expanded m x m matrix is staged in a
 * fixed per-device area of 32 MiB (33554432 bytes) that the library holds until the process ends, so
 * m * m * element size must not exceed 33554432 (else it returns non-zero and says why).  The stream
 * must belong to the current device.  Calls on one device share that area: calls that overlap on one
 * device (two streams, or two host threads) race and must be ordered by the caller. */
int bhblas_symm_noalloc(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb,
                        void *C, int64_t ldc, void *stream);
/* C = A B, A m x m Hermitian given by its upper triangle; complex only (cblas_?hemm, left, upper). */
int bhblas_hemm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                int64_t ldc, void *stream);
/* B := A B in place, A m x m unit upper triangular, B m x n (cblas_?trmm, left, upper, no-trans, unit). */
int bhblas_trmm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *B, int64_t ldb, void *stream);
/* B := A^-1 B in place (solves A X = B), A m x m unit upper triangular (cblas_?trsm, same options). */
int bhblas_trsm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *B, int64_t ldb, void *stream);
const char *bhblas_last_error(void);

/* Extension methods blas_<name> (templates/methods.json "methods"). */
bh_extmethod_impl *blas_gemm_create(void);
void blas_gemm_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_gemmt_create(void);
void blas_gemmt_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_symm_create(void);
void blas_symm_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_hemm_create(void);
void blas_hemm_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_syrk_create(void);
void blas_syrk_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_herk_create(void);
void blas_herk_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_syr2k_create(void);
void blas_syr2k_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_her2k_create(void);
void blas_her2k_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_trmm_create(void);
void blas_trmm_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_trsm_create(void);
void blas_trsm_destroy(bh_extmethod_impl *self);

#ifdef __cplusplus
}
#endif

#endif
