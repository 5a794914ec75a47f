/* bh_blas_hip — device BLAS extension methods for the hip vector engine.
 *
 * libbh_blas_hip.so is an extension-method library in the reference's sense
 * (include/bohrium/bh_extmethod.hpp:57-125, found by core/bh_extmethod.cpp:33-90
 * through the engine's `libs` option): for each method it exports
 * `<name>_create` / `<name>_destroy` returning an ExtmethodImpl.  It replaces
 * the reference's CBLAS-backed extmethods/blas (templates/body.tpl:1-50,
 * body_func.tpl:1-30, methods.json) for the methods below, on the device:
 *
 *   blas_gemm    C := A B      A m x k, B k x n, C m x n   (body_func.tpl: alpha 1, beta 0)
 *   blas_gemmt   C := A^T B    A m x k, B m x n, C k x n
 *   blas_syrk    C := A A^T    upper triangle of C (m x m), the strictly lower part is left as it was
 *
 * Instruction operands are (C, A, B) for gemm/gemmt and (C, A) for syrk, as the
 * reference front end queues them (bridge/npbackend/bohrium/blas.py:12-45); the
 * front end applies alpha and beta before the call.  Matrices are row-major and
 * C-contiguous; float32 and float64 run on MFMA (v_mfma_f32_16x16x4f32,
 * v_mfma_f64_16x16x4f64), complex64/complex128 on a tiled FMA kernel.
 *
 * bhblas_gemm is the same device routine as a plain C entry point (device
 * pointers, the engine stream or any HIP stream) for callers outside a
 * component stack and for benchmarks.
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
/* The upper triangle (i <= j) of C = A * A^T, A m x k, C m x m. */
int bhblas_syrk(int dtype, int64_t m, int64_t k, const void *A, int64_t lda, void *C, int64_t ldc, void *stream);
const char *bhblas_last_error(void);

bh_extmethod_impl *blas_gemm_create(void);
void blas_gemm_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_gemmt_create(void);
void blas_gemmt_destroy(bh_extmethod_impl *self);
bh_extmethod_impl *blas_syrk_create(void);
void blas_syrk_destroy(bh_extmethod_impl *self);

#ifdef __cplusplus
}
#endif

#endif
