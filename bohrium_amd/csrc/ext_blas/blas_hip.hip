/* Device BLAS for the hip engine: GEMM on the gfx950 matrix cores, exposed as
 * the reference's blas_* extension methods (include/bh_blas_hip.h).
 *
 * Real types: one workgroup of 4 waves computes a BM x BN tile of C; wave w owns
 * the (w / 2, w % 2) quarter as TM x TN MFMA tiles of 16 x 16.  Per K step (BK
 * of 16 or 32) the workgroup stages op(A) (BM x BK) and op(B) (BK x BN) in LDS
 * with padded rows so that the 16-lane groups of one MFMA operand read fall in
 * disjoint banks; the LDS tiles are double-buffered and the next K step is
 * fetched into registers while the current one runs on the matrix cores.  MFMA operand layout (16 x 16 x 4, measured on gfx950): lane l
 * supplies A[l % 16][l / 16] and B[l / 16][l % 16]; its 4 results are
 * D[4 (l / 16) + r][l % 16] for f32 and D[4 r + l / 16][l % 16] for f64, r = 0..3.
 *
 * Complex types use a 16 x 16 LDS-tiled kernel with complex fused multiply-adds.
 */
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include <bohrium/bh_extmethod.hpp>
#include <bohrium/bh_instruction.hpp>

#include "../../../include/bh_blas_hip.h"
#include "../../../include/bh_ve_hip.h"

namespace {

typedef double bh_d4 __attribute__((ext_vector_type(4)));
typedef float bh_f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef bh_d4 Acc;
    static __device__ __forceinline__ Acc step(double a, double b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the 16 x 16 tile that result r of `lane` belongs to
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef bh_f4 Acc;
    static __device__ __forceinline__ Acc step(float a, float b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kThreads = 256;
constexpr int kPad = 16;

template <typename T> struct Vec;
template <> struct Vec<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int n = 2;
};
template <> struct Vec<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int n = 4;
};

// One operand tile: ROWS rows (m of op(A) or n of op(B)) by kBK steps of k (16 or 32).  KCONT: k is the
// contiguous dimension in memory (element (r, k) at P[r * ld + k]); otherwise r is
// (element at P[k * ld + r]).  A thread moves chunks of V elements along the contiguous dimension
// (one 16-byte load on interior tiles).  LDS layout: [ROWS][kBK + KPAD] when KCONT, else
// [kBK][ROWS + kPad]; with these pads the 16-lane groups of an MFMA operand read (16 rows at one k,
// the next group at k + 1) and the chunk stores fall in disjoint banks.
template <typename T, bool KCONT, int ROWS, int kBK>
struct Tile {
    static constexpr int V = Vec<T>::n;
    static constexpr int PER = ROWS * kBK / kThreads;
    static constexpr int CH = PER / V;
    static constexpr int KPAD = sizeof(T) == 8 ? 2 : 4;
    static constexpr int LDS = KCONT ? ROWS * (kBK + KPAD) : kBK * (ROWS + kPad);
    static_assert(CH * V == PER && CH * kThreads * V == ROWS * kBK, "tile / thread count");

    static __device__ __forceinline__ void chunk(int c, int &r, int &k) {
        if (KCONT) {
            r = c / (kBK / V);
            k = (c % (kBK / V)) * V;
        } else {
            k = c / (ROWS / V);
            r = (c % (ROWS / V)) * V;
        }
    }
    // vec: the whole tile is in bounds and every chunk is 16-byte aligned.
    static __device__ __forceinline__ void fetch(const T *__restrict__ P, int64_t ld, int64_t r0, int64_t k0, int64_t R,
                                                 int64_t K, bool vec, T *reg) {
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            int r, k;
            chunk(static_cast<int>(threadIdx.x) + e * kThreads, r, k);
            const int64_t gr = r0 + r, gk = k0 + k;
            if (vec) {
                const typename Vec<T>::type v =
                    *reinterpret_cast<const typename Vec<T>::type *>(KCONT ? P + gr * ld + gk : P + gk * ld + gr);
#pragma unroll
                for (int j = 0; j < V; ++j) reg[e * V + j] = v[j];
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const int64_t rr = KCONT ? gr : gr + j, kk = KCONT ? gk + j : gk;
                    reg[e * V + j] = (rr < R && kk < K) ? P[KCONT ? rr * ld + kk : kk * ld + rr] : T(0);
                }
            }
        }
    }
    static __device__ __forceinline__ void store(T *lds, const T *reg) {
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            int r, k;
            chunk(static_cast<int>(threadIdx.x) + e * kThreads, r, k);
            typename Vec<T>::type v;
#pragma unroll
            for (int j = 0; j < V; ++j) v[j] = reg[e * V + j];
            *reinterpret_cast<typename Vec<T>::type *>(KCONT ? lds + r * (kBK + KPAD) + k : lds + k * (ROWS + kPad) + r) = v;
        }
    }
    static __device__ __forceinline__ T read(const T *lds, int r, int k) {
        return KCONT ? lds[r * (kBK + KPAD) + k] : lds[k * (ROWS + kPad) + r];
    }
};

// C = op(A) op(B); op(A)(m, k) = TA ? A[k * lda + m] : A[m * lda + k],
// op(B)(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n].  UPPER: only C(m, n) with m <= n is stored.
// vec_ok: lda and ldb are multiples of the 16-byte vector and A, B are 16-byte aligned.
// The LDS tiles are double-buffered: while the matrix cores work on one K step, the next one is
// loaded into registers and then stored into the other buffer; one barrier per K step.
template <typename T, bool TA, bool TB, bool UPPER, int TM, int TN, int BK>
__global__ void __launch_bounds__(kThreads)
gemm_mfma(const T *__restrict__ A, const T *__restrict__ B, T *__restrict__ C, int64_t M, int64_t N, int64_t K,
          int64_t lda, int64_t ldb, int64_t ldc, int vec_ok) {
    constexpr int BM = 32 * TM, BN = 32 * TN;
    typedef Tile<T, !TA, BM, BK> ATile;
    typedef Tile<T, TB, BN, BK> BTile;
    __shared__ __attribute__((aligned(16))) T As[2][ATile::LDS];
    __shared__ __attribute__((aligned(16))) T Bs[2][BTile::LDS];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t bm = static_cast<int64_t>(blockIdx.y) * BM, bn = static_cast<int64_t>(blockIdx.x) * BN;
    if (UPPER && bm > bn + BN - 1) return;  // the whole tile lies below the diagonal
    const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
    const bool a_in = vec_ok && bm + BM <= M, b_in = vec_ok && bn + BN <= N;

    typename Mfma<T>::Acc acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);

    T ra[ATile::PER], rb[BTile::PER];
    auto fetch = [&](int64_t k0) {
        const bool kin = k0 + BK <= K;
        ATile::fetch(A, lda, bm, k0, M, K, a_in && kin, ra);
        BTile::fetch(B, ldb, bn, k0, N, K, b_in && kin, rb);
    };

    fetch(0);
    ATile::store(As[0], ra);
    BTile::store(Bs[0], rb);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) fetch(k0 + BK);
        const T *as = As[buf], *bs = Bs[buf];
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const int kr = ks + (lane >> 4);
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = ATile::read(as, wm + i * 16 + (lane & 15), kr);
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = BTile::read(bs, wn + j * 16 + (lane & 15), kr);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::step(a[i], b[j], acc[i][j]);
        }
        if (more) {
            ATile::store(As[buf ^ 1], ra);
            BTile::store(Bs[buf ^ 1], rb);
        }
        // the other buffer is complete, and every wave is done reading this one before it is refilled
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = bm + wm + i * 16 + Mfma<T>::row(lane, r);
                const int64_t col = bn + wn + j * 16 + (lane & 15);
                if (row < M && col < N && (!UPPER || row <= col)) C[row * ldc + col] = acc[i][j][r];
            }
}

// Complex: 16 x 16 threads, each one element of C; K in steps of 16 through LDS.
template <typename R> struct Cplx {
    R re, im;
};

template <typename R, bool TA, bool TB, bool UPPER>
__global__ void __launch_bounds__(256)
gemm_complex(const Cplx<R> *__restrict__ A, const Cplx<R> *__restrict__ B, Cplx<R> *__restrict__ C, int64_t M,
             int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc) {
    __shared__ Cplx<R> As[16][17];
    __shared__ Cplx<R> Bs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t row = static_cast<int64_t>(blockIdx.y) * 16 + ty, col = static_cast<int64_t>(blockIdx.x) * 16 + tx;
    R sr = 0, si = 0;
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
        const int64_t ka = k0 + tx, kb = k0 + ty;
        Cplx<R> z = {0, 0};
        As[ty][tx] = (row < M && ka < K) ? (TA ? A[ka * lda + row] : A[row * lda + ka]) : z;
        Bs[ty][tx] = (kb < K && col < N) ? (TB ? B[col * ldb + kb] : B[kb * ldb + col]) : z;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const Cplx<R> a = As[ty][k], b = Bs[k][tx];
            sr = fma(a.re, b.re, sr);
            sr = fma(-a.im, b.im, sr);
            si = fma(a.re, b.im, si);
            si = fma(a.im, b.re, si);
        }
        __syncthreads();
    }
    if (row < M && col < N && (!UPPER || row <= col)) C[row * ldc + col] = {sr, si};
}

thread_local std::string g_err;

template <typename T, bool TA, bool TB, bool UPPER>
void launch_real(int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                 int64_t ldc, hipStream_t s) {
    const int64_t v = Vec<T>::n;
    const int vec_ok = lda % v == 0 && ldb % v == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(B) % 16 == 0;
    // 128 x 128 tiles when the matrix fills the device with them, 64 x 64 otherwise; K steps of
    // BH_BLAS_BK elements (16 or 32; 16 measured faster on MI355X: 8192^3 fp64 59.5 vs 40.6,
    // fp32 129.1 vs 120.0 TFLOP/s).
    static const int bk = [] {
        const char *e = std::getenv("BH_BLAS_BK");
        return e != nullptr && std::atoi(e) == 32 ? 32 : 16;
    }();
    const bool big = ((m + 127) / 128) * ((n + 127) / 128) >= 512;
    const T *a = static_cast<const T *>(A), *b = static_cast<const T *>(B);
    T *c = static_cast<T *>(C);
    if (big) {
        const dim3 grid(static_cast<unsigned>((n + 127) / 128), static_cast<unsigned>((m + 127) / 128));
        if (bk == 32)
            gemm_mfma<T, TA, TB, UPPER, 4, 4, 32><<<grid, kThreads, 0, s>>>(a, b, c, m, n, k, lda, ldb, ldc, vec_ok);
        else
            gemm_mfma<T, TA, TB, UPPER, 4, 4, 16><<<grid, kThreads, 0, s>>>(a, b, c, m, n, k, lda, ldb, ldc, vec_ok);
    } else {
        const dim3 grid(static_cast<unsigned>((n + 63) / 64), static_cast<unsigned>((m + 63) / 64));
        if (bk == 32)
            gemm_mfma<T, TA, TB, UPPER, 2, 2, 32><<<grid, kThreads, 0, s>>>(a, b, c, m, n, k, lda, ldb, ldc, vec_ok);
        else
            gemm_mfma<T, TA, TB, UPPER, 2, 2, 16><<<grid, kThreads, 0, s>>>(a, b, c, m, n, k, lda, ldb, ldc, vec_ok);
    }
}

template <typename R, bool TA, bool TB, bool UPPER>
void launch_complex(int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                    int64_t ldc, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>((n + 15) / 16), static_cast<unsigned>((m + 15) / 16));
    gemm_complex<R, TA, TB, UPPER><<<grid, 256, 0, s>>>(static_cast<const Cplx<R> *>(A), static_cast<const Cplx<R> *>(B),
                                                        static_cast<Cplx<R> *>(C), m, n, k, lda, ldb, ldc);
}

template <bool TA, bool TB, bool UPPER>
int dispatch(int dtype, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb,
             void *C, int64_t ldc, void *stream) {
    if (m < 0 || n < 0 || k < 0) {
        g_err = "bhblas: negative matrix size";
        return 1;
    }
    // grid.y carries the tile rows (at most 65535 * 16 rows for the complex kernel)
    if ((m + 15) / 16 > 65535 || (n + 15) / 16 > 0x7fffffffLL) {
        g_err = "bhblas: matrix too large for one launch";
        return 1;
    }
    if (m == 0 || n == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (dtype) {
        case BHBLAS_F32: launch_real<float, TA, TB, UPPER>(m, n, k, A, lda, B, ldb, C, ldc, s); break;
        case BHBLAS_F64: launch_real<double, TA, TB, UPPER>(m, n, k, A, lda, B, ldb, C, ldc, s); break;
        case BHBLAS_C64: launch_complex<float, TA, TB, UPPER>(m, n, k, A, lda, B, ldb, C, ldc, s); break;
        case BHBLAS_C128: launch_complex<double, TA, TB, UPPER>(m, n, k, A, lda, B, ldb, C, ldc, s); break;
        default: g_err = "bhblas: unsupported type"; return 1;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_err = std::string("bhblas: launch failed: ") + hipGetErrorString(e);
        return 1;
    }
    return 0;
}

// ------------------------------------------------------------ extension methods

using bohrium::extmethod::ExtmethodImpl;

int blas_type(const bh_view &v) {
    switch (v.base->dtype()) {
        case bh_type::FLOAT32: return BHBLAS_F32;
        case bh_type::FLOAT64: return BHBLAS_F64;
        case bh_type::COMPLEX64: return BHBLAS_C64;
        case bh_type::COMPLEX128: return BHBLAS_C128;
        default: return -1;
    }
}

// A 2-D C-contiguous view: row-major with unit inner stride and outer stride = columns.
void require_matrix(const bh_view &v, const char *what, const std::string &method) {
    if (v.isConstant() || v.ndim != 2 || v.stride[1] != 1 || v.stride[0] != v.shape[1])
        throw std::runtime_error(method + ": " + what + " must be a C-contiguous 2-d array");
}

enum class Kind { GEMM, GEMMT, SYRK };

class BlasMethod : public ExtmethodImpl {
    Kind kind;
    std::string name;

public:
    BlasMethod(Kind k, const char *n) : kind(k), name(n) {}
    void execute(bh_instruction *instr, void *arg) override {
        const bh_hip_ext_context *ctx = static_cast<const bh_hip_ext_context *>(arg);
        if (ctx == nullptr || ctx->version != 1) throw std::runtime_error(name + ": needs the hip engine's context");
        const size_t nops = kind == Kind::SYRK ? 2 : 3;
        if (instr->operand.size() != nops) throw std::runtime_error(name + ": wrong number of operands");
        const bh_view &C = instr->operand[0], &A = instr->operand[1];
        require_matrix(C, "C", name);
        require_matrix(A, "A", name);
        const int dt = blas_type(A);
        if (dt < 0) throw std::runtime_error(name + ": " + bh_type_text(A.base->dtype()) + " not supported by BLAS");
        if (C.base->dtype() != A.base->dtype()) throw std::runtime_error(name + ": operands differ in type");
        const int64_t am = A.shape[0], ak = A.shape[1];
        const int esize = bh_type_size(A.base->dtype());
        auto ptr = [&](const bh_view &v) {
            return static_cast<char *>(ctx->device_pointer(ctx->engine, v.base)) + v.start * esize;
        };
        int rc = 0;
        if (kind == Kind::SYRK) {
            if (C.shape[0] != am || C.shape[1] != am) throw std::runtime_error(name + ": C must be m x m for A m x k");
            if (ctx->dry_run) return;
            rc = bhblas_syrk(dt, am, ak, ptr(A), ak, ptr(C), am, ctx->stream);
        } else {
            const bh_view &B = instr->operand[2];
            require_matrix(B, "B", name);
            if (B.base->dtype() != A.base->dtype()) throw std::runtime_error(name + ": operands differ in type");
            if (kind == Kind::GEMM) {
                if (B.shape[0] != ak || C.shape[0] != am || C.shape[1] != B.shape[1])
                    throw std::runtime_error(name + ": shapes do not chain (A m x k, B k x n, C m x n)");
                if (ctx->dry_run) return;
                rc = bhblas_gemm(dt, 0, am, B.shape[1], ak, ptr(A), ak, ptr(B), B.shape[1], ptr(C), C.shape[1],
                                 ctx->stream);
            } else {
                if (B.shape[0] != am || C.shape[0] != ak || C.shape[1] != B.shape[1])
                    throw std::runtime_error(name + ": shapes do not chain (A m x k, B m x n, C k x n)");
                if (ctx->dry_run) return;
                rc = bhblas_gemm(dt, 1, ak, B.shape[1], am, ptr(A), ak, ptr(B), B.shape[1], ptr(C), C.shape[1],
                                 ctx->stream);
            }
        }
        if (rc != 0) throw std::runtime_error(name + ": " + g_err);
    }
};

} // namespace

extern "C" {

int bhblas_gemm(int dtype, int trans_a, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B,
                int64_t ldb, void *C, int64_t ldc, void *stream) {
    return trans_a ? dispatch<true, false, false>(dtype, m, n, k, A, lda, B, ldb, C, ldc, stream)
                   : dispatch<false, false, false>(dtype, m, n, k, A, lda, B, ldb, C, ldc, stream);
}

int bhblas_syrk(int dtype, int64_t m, int64_t k, const void *A, int64_t lda, void *C, int64_t ldc, void *stream) {
    // C = A A^T: op(B)(kk, n) = A[n][kk], i.e. B = A read transposed.
    return dispatch<false, true, true>(dtype, m, m, k, A, lda, A, lda, C, ldc, stream);
}

const char *bhblas_last_error(void) { return g_err.c_str(); }

bh_extmethod_impl *blas_gemm_create(void) { return new BlasMethod(Kind::GEMM, "blas_gemm"); }
void blas_gemm_destroy(bh_extmethod_impl *self) { delete self; }
bh_extmethod_impl *blas_gemmt_create(void) { return new BlasMethod(Kind::GEMMT, "blas_gemmt"); }
void blas_gemmt_destroy(bh_extmethod_impl *self) { delete self; }
bh_extmethod_impl *blas_syrk_create(void) { return new BlasMethod(Kind::SYRK, "blas_syrk"); }
void blas_syrk_destroy(bh_extmethod_impl *self) { delete self; }

} // extern "C"
