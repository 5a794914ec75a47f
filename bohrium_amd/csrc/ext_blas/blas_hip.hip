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

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <mutex>
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

// C = op(A) op(B), or alpha op(A) op(B) + beta C when SCALED (C is not read when beta is 0); op(A)(m, k)
// = TA ? A[k * lda + m] : A[m * lda + k], op(B)(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n].  UPPER:
// only C(m, n) with m <= n is stored.  The scaled epilogue is a separate instantiation: in the plain
// product it costs the f64 128 x 128 kernel ~120 VGPRs and half its throughput.
// vec_ok: lda and ldb are multiples of the 16-byte vector and A, B are 16-byte aligned.
// The LDS tiles are double-buffered: while the matrix cores work on one K step, the next one is
// loaded into registers and then stored into the other buffer; one barrier per K step.
template <typename T, bool TA, bool TB, bool UPPER, int TM, int TN, int BK, bool SCALED>
__global__ void __launch_bounds__(kThreads)
gemm_mfma(const T *__restrict__ A, const T *__restrict__ B, T *__restrict__ C, int64_t M, int64_t N, int64_t K,
          int64_t lda, int64_t ldb, int64_t ldc, int vec_ok, T alpha, T beta) {
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
                if (row < M && col < N && (!UPPER || row <= col)) {
                    if (SCALED) {
                        T v = alpha * acc[i][j][r];
                        if (beta != T(0)) v += beta * C[row * ldc + col];
                        C[row * ldc + col] = v;
                    } else {
                        C[row * ldc + col] = acc[i][j][r];
                    }
                }
            }
}

// Complex: 16 x 16 threads, each one element of C; K in steps of 16 through LDS.
template <typename R> struct Cplx {
    R re, im;
};

template <typename T> __device__ __forceinline__ T conj_of(T x) { return x; }
template <typename R> __device__ __forceinline__ Cplx<R> conj_of(Cplx<R> z) { return {z.re, -z.im}; }
template <typename T> __device__ __forceinline__ T real_of(T x) { return x; }
template <typename R> __device__ __forceinline__ Cplx<R> real_of(Cplx<R> z) { return {z.re, R(0)}; }
template <typename T> struct One {
    static __device__ __forceinline__ T get() { return T(1); }
};
template <typename R> struct One<Cplx<R>> {
    static __device__ __forceinline__ Cplx<R> get() { return {R(1), R(0)}; }
};
// s - a x
template <typename T> __device__ __forceinline__ T msub(T s, T a, T x) { return fma(-a, x, s); }
template <typename R> __device__ __forceinline__ Cplx<R> msub(Cplx<R> s, Cplx<R> a, Cplx<R> x) {
    return {fma(a.im, x.im, fma(-a.re, x.re, s.re)), fma(-a.im, x.re, fma(-a.re, x.im, s.im))};
}

// C = alpha op(A) op(B) + beta C with op(B)(k, n) = conj(B[n * ldb + k]) when CONJB; REALDIAG stores
// the diagonal with a zero imaginary part (Hermitian results).
template <typename R, bool TA, bool TB, bool CONJB, bool UPPER, bool REALDIAG>
__global__ void __launch_bounds__(256)
gemm_complex(const Cplx<R> *__restrict__ A, const Cplx<R> *__restrict__ B, Cplx<R> *__restrict__ C, int64_t M,
             int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc, R alpha, R beta) {
    __shared__ Cplx<R> As[16][17];
    __shared__ Cplx<R> Bs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t row = static_cast<int64_t>(blockIdx.y) * 16 + ty, col = static_cast<int64_t>(blockIdx.x) * 16 + tx;
    if (UPPER && static_cast<int64_t>(blockIdx.y) > static_cast<int64_t>(blockIdx.x)) return;  // below the diagonal
    R sr = 0, si = 0;
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
        const int64_t ka = k0 + tx, kb = k0 + ty;
        Cplx<R> z = {0, 0};
        As[ty][tx] = (row < M && ka < K) ? (TA ? A[ka * lda + row] : A[row * lda + ka]) : z;
        if (kb < K && col < N) {
            const Cplx<R> b = TB ? B[col * ldb + kb] : B[kb * ldb + col];
            Bs[ty][tx] = CONJB ? conj_of(b) : b;
        } else {
            Bs[ty][tx] = z;
        }
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
    if (row < M && col < N && (!UPPER || row <= col)) {
        Cplx<R> v = {alpha * sr, alpha * si};
        if (beta != R(0)) {
            const Cplx<R> c = C[row * ldc + col];
            v.re = fma(beta, c.re, v.re);
            v.im = fma(beta, c.im, v.im);
        }
        if (REALDIAG && row == col) v.im = 0;
        C[row * ldc + col] = v;
    }
}

// out (m x m, leading dimension m) from the upper triangle of A (i <= j read, the rest not
// referenced): SYM mirrors it, HERM mirrors it conjugated with a real diagonal, UNIT is the unit
// upper triangular matrix (ones on the diagonal, zeros below).
enum class Fill { SYM, HERM, UNIT };

template <typename T, Fill F>
__global__ void __launch_bounds__(256) fill_from_upper(const T *__restrict__ A, int64_t lda, T *__restrict__ out,
                                                       int64_t m) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 16 + (threadIdx.x & 15);
    const int64_t i = static_cast<int64_t>(blockIdx.y) * 16 + (threadIdx.x >> 4);
    if (i >= m || j >= m) return;
    T v;
    if (i < j)
        v = A[i * lda + j];
    else if (i > j)
        v = F == Fill::UNIT ? T{} : F == Fill::HERM ? conj_of(A[j * lda + i]) : A[j * lda + i];
    else
        v = F == Fill::UNIT ? One<T>::get() : F == Fill::HERM ? real_of(A[i * lda + i]) : A[i * lda + i];
    out[i * m + j] = v;
}

// Solves U X = B in place for one diagonal block of a unit upper triangular system: U is the nb x nb
// block at A (leading dimension lda; its diagonal and lower part are not referenced), B the nb rows
// at B (leading dimension ldb, n columns).  One thread per column keeps the column in registers
// (rows past nb are zero and not stored) and substitutes backwards with U's block from LDS.
template <typename T, int NB>
__global__ void __launch_bounds__(256) trsm_unit_upper_block(const T *__restrict__ A, int64_t lda,
                                                             T *__restrict__ B, int64_t ldb, int nb, int64_t n) {
    __shared__ T U[NB][NB + 1];
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int i = e / NB, k = e % NB;
        U[i][k] = (k > i && k < nb) ? A[static_cast<int64_t>(i) * lda + k] : T{};
    }
    __syncthreads();
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    T x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = i < nb ? B[i * ldb + j] : T{};
#pragma unroll
    for (int i = NB - 2; i >= 0; --i) {
        T s = x[i];
#pragma unroll
        for (int k = i + 1; k < NB; ++k) s = msub(s, U[i][k], x[k]);
        x[i] = s;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i < nb) B[i * ldb + j] = x[i];
}

thread_local std::string g_err;

// One GEMM call: C = alpha op(A) op(B) + beta C.
struct Gemm {
    int dtype;
    int64_t m, n, k;
    const void *A;
    int64_t lda;
    const void *B;
    int64_t ldb;
    void *C;
    int64_t ldc;
    double alpha, beta;
    hipStream_t s;
};

template <typename T, bool TA, bool TB, bool UPPER, bool SCALED>
void launch_real_tiles(const Gemm &g, int vec_ok) {
    // 128 x 128 tiles when the matrix fills the device with them, 64 x 64 otherwise; K steps of
    // BH_BLAS_BK elements (16 or 32; 16 measured faster on MI355X: 8192^3 fp64 59.5 vs 40.6,
    // fp32 129.1 vs 120.0 TFLOP/s).
    static const int bk = [] {
        const char *e = std::getenv("BH_BLAS_BK");
        return e != nullptr && std::atoi(e) == 32 ? 32 : 16;
    }();
    const int64_t m = g.m, n = g.n;
    const bool big = ((m + 127) / 128) * ((n + 127) / 128) >= 512;
    const T *a = static_cast<const T *>(g.A), *b = static_cast<const T *>(g.B);
    T *c = static_cast<T *>(g.C);
    const T al = static_cast<T>(g.alpha), be = static_cast<T>(g.beta);
    if (big) {
        const dim3 grid(static_cast<unsigned>((n + 127) / 128), static_cast<unsigned>((m + 127) / 128));
        if (bk == 32)
            gemm_mfma<T, TA, TB, UPPER, 4, 4, 32, SCALED>
                <<<grid, kThreads, 0, g.s>>>(a, b, c, m, n, g.k, g.lda, g.ldb, g.ldc, vec_ok, al, be);
        else
            gemm_mfma<T, TA, TB, UPPER, 4, 4, 16, SCALED>
                <<<grid, kThreads, 0, g.s>>>(a, b, c, m, n, g.k, g.lda, g.ldb, g.ldc, vec_ok, al, be);
    } else {
        const dim3 grid(static_cast<unsigned>((n + 63) / 64), static_cast<unsigned>((m + 63) / 64));
        if (bk == 32)
            gemm_mfma<T, TA, TB, UPPER, 2, 2, 32, SCALED>
                <<<grid, kThreads, 0, g.s>>>(a, b, c, m, n, g.k, g.lda, g.ldb, g.ldc, vec_ok, al, be);
        else
            gemm_mfma<T, TA, TB, UPPER, 2, 2, 16, SCALED>
                <<<grid, kThreads, 0, g.s>>>(a, b, c, m, n, g.k, g.lda, g.ldb, g.ldc, vec_ok, al, be);
    }
}

template <typename T, bool TA, bool TB, bool UPPER>
bool launch_real(const Gemm &g) {
    const int64_t v = Vec<T>::n;
    const int vec_ok = g.lda % v == 0 && g.ldb % v == 0 && reinterpret_cast<uintptr_t>(g.A) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(g.B) % 16 == 0;
    if (g.alpha == 1.0 && g.beta == 0.0) {
        launch_real_tiles<T, TA, TB, UPPER, false>(g, vec_ok);
    } else if constexpr (!TA) {  // trsm's updates and syr2k's second pass
        launch_real_tiles<T, TA, TB, UPPER, true>(g, vec_ok);
    } else {
        g_err = "bhblas: scaled product with a transposed A is not provided";
        return false;
    }
    return true;
}

template <typename R, bool TA, bool TB, bool CONJB, bool UPPER, bool REALDIAG>
void launch_complex(const Gemm &g) {
    const dim3 grid(static_cast<unsigned>((g.n + 15) / 16), static_cast<unsigned>((g.m + 15) / 16));
    gemm_complex<R, TA, TB, CONJB, UPPER, REALDIAG><<<grid, 256, 0, g.s>>>(
        static_cast<const Cplx<R> *>(g.A), static_cast<const Cplx<R> *>(g.B), static_cast<Cplx<R> *>(g.C), g.m, g.n,
        g.k, g.lda, g.ldb, g.ldc, static_cast<R>(g.alpha), static_cast<R>(g.beta));
}

int launch_status() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_err = std::string("bhblas: launch failed: ") + hipGetErrorString(e);
        return 1;
    }
    return 0;
}

bool is_complex(int dtype) { return dtype == BHBLAS_C64 || dtype == BHBLAS_C128; }

int64_t elem_size(int dtype) {
    switch (dtype) {
        case BHBLAS_F32: return 4;
        case BHBLAS_F64: return 8;
        case BHBLAS_C64: return 8;
        case BHBLAS_C128: return 16;
        default: return 0;
    }
}

// Checks the sizes and the type of a call over m x n (and k); false with g_err set when they are
// unusable.  Grid y carries 16-row tiles, at most 65535 of them.
bool check_call(int dtype, int64_t m, int64_t n, int64_t k) {
    if (elem_size(dtype) == 0) {
        g_err = "bhblas: unsupported type";
        return false;
    }
    if (m < 0 || n < 0 || k < 0) {
        g_err = "bhblas: negative matrix size";
        return false;
    }
    if ((m + 15) / 16 > 65535 || (n + 15) / 16 > 0x7fffffffLL) {
        g_err = "bhblas: matrix too large for one launch";
        return false;
    }
    return true;
}

template <bool TA, bool TB, bool UPPER, bool CONJB = false, bool REALDIAG = false>
int dispatch(const Gemm &g) {
    if (!check_call(g.dtype, g.m, g.n, g.k)) return 1;
    if (g.m == 0 || g.n == 0) return 0;
    switch (g.dtype) {
        case BHBLAS_F32:
            if (!launch_real<float, TA, TB, UPPER>(g)) return 1;
            break;
        case BHBLAS_F64:
            if (!launch_real<double, TA, TB, UPPER>(g)) return 1;
            break;
        case BHBLAS_C64: launch_complex<float, TA, TB, CONJB, UPPER, REALDIAG>(g); break;
        case BHBLAS_C128: launch_complex<double, TA, TB, CONJB, UPPER, REALDIAG>(g); break;
    }
    return launch_status();
}

Gemm make_gemm(int dtype, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb,
               void *C, int64_t ldc, void *stream, double alpha = 1.0, double beta = 0.0) {
    return Gemm{dtype, m, n, k, A, lda, B, ldb, C, ldc, alpha, beta, static_cast<hipStream_t>(stream)};
}

// Stream-ordered device temporary, released on the same stream.
struct DeviceTemp {
    void *p = nullptr;
    hipStream_t s;
    explicit DeviceTemp(hipStream_t st) : s(st) {}
    bool alloc(size_t bytes) {
        if (bytes == 0) return true;
        if (hipMallocAsync(&p, bytes, s) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            g_err = "bhblas: out of device memory for a temporary of " + std::to_string(bytes) + " bytes";
            return false;
        }
        return true;
    }
    ~DeviceTemp() {
        if (p != nullptr) (void)hipFreeAsync(p, s);
    }
};

template <Fill F> int fill(int dtype, const void *A, int64_t lda, void *out, int64_t m, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>((m + 15) / 16), static_cast<unsigned>((m + 15) / 16));
    switch (dtype) {
        case BHBLAS_F32: fill_from_upper<float, F><<<grid, 256, 0, s>>>(static_cast<const float *>(A), lda,
                                                                         static_cast<float *>(out), m); break;
        case BHBLAS_F64: fill_from_upper<double, F><<<grid, 256, 0, s>>>(static_cast<const double *>(A), lda,
                                                                          static_cast<double *>(out), m); break;
        case BHBLAS_C64: fill_from_upper<Cplx<float>, F><<<grid, 256, 0, s>>>(
            static_cast<const Cplx<float> *>(A), lda, static_cast<Cplx<float> *>(out), m); break;
        case BHBLAS_C128: fill_from_upper<Cplx<double>, F><<<grid, 256, 0, s>>>(
            static_cast<const Cplx<double> *>(A), lda, static_cast<Cplx<double> *>(out), m); break;
    }
    return launch_status();
}

// C = F(A) B with F(A) expanded from A's upper triangle into an m x m buffer: a stream-ordered
// temporary, or `staging` when the caller supplies one of at least m * m elements.
template <Fill F>
int expanded_product(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb,
                     void *C, int64_t ldc, void *stream, void *staging = nullptr) {
    if (!check_call(dtype, m, n, m)) return 1;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DeviceTemp full(s);
    void *expanded = staging;
    if (expanded == nullptr) {
        if (!full.alloc(static_cast<size_t>(m * m * elem_size(dtype)))) return 1;
        expanded = full.p;
    }
    if (fill<F>(dtype, A, lda, expanded, m, s) != 0) return 1;
    return dispatch<false, false, false>(make_gemm(dtype, m, n, m, expanded, m, B, ldb, C, ldc, stream));
}

template <typename T, int NB>
void trsm_block(const void *A, int64_t lda, void *B, int64_t ldb, int nb, int64_t n, hipStream_t s) {
    trsm_unit_upper_block<T, NB><<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(
        static_cast<const T *>(A), lda, static_cast<T *>(B), ldb, nb, n);
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

enum class Kind { GEMM, GEMMT, SYMM, HEMM, SYRK, HERK, SYR2K, HER2K, TRMM, TRSM };

// The reference's methods table (extmethods/blas/templates/methods.json): operands (C, A[, B]),
// row-major, CblasUpper / CblasLeft / CblasNoTrans / CblasUnit, alpha 1 and beta 0.
class BlasMethod : public ExtmethodImpl {
    Kind kind;
    std::string name;

public:
    BlasMethod(Kind k, const char *n) : kind(k), name(n) {}
    void execute(bh_instruction *instr, void *arg) override {
        const bh_hip_ext_context *ctx = static_cast<const bh_hip_ext_context *>(arg);
        if (ctx == nullptr || ctx->version != 1) throw std::runtime_error(name + ": needs the hip engine's context");
        const bool only_a = kind == Kind::SYRK || kind == Kind::HERK;
        if (instr->operand.size() != (only_a ? 2u : 3u)) throw std::runtime_error(name + ": wrong number of operands");
        const bool triangular = kind == Kind::TRMM || kind == Kind::TRSM;  // C is not used
        const bh_view &C = instr->operand[0], &A = instr->operand[1];
        if (!triangular) require_matrix(C, "C", name);
        require_matrix(A, "A", name);
        const int dt = blas_type(A);
        const bool complex_only = kind == Kind::HEMM || kind == Kind::HERK || kind == Kind::HER2K;
        if (dt < 0 || (complex_only && !is_complex(dt)))
            throw std::runtime_error(std::string(bh_type_text(A.base->dtype())) + " not supported by BLAS for '" +
                                     name + "'.");
        if (!triangular && C.base->dtype() != A.base->dtype())
            throw std::runtime_error(name + ": operands differ in type");
        const int64_t am = A.shape[0], ak = A.shape[1];
        const int esize = bh_type_size(A.base->dtype());
        auto ptr = [&](const bh_view &v) {
            return static_cast<char *>(ctx->device_pointer(ctx->engine, v.base)) + v.start * esize;
        };
        auto shape_error = [&](const char *what) { throw std::runtime_error(name + ": " + what); };
        if (only_a) {
            if (C.shape[0] != am || C.shape[1] != am) shape_error("C must be n x n for A n x k");
            if (ctx->dry_run) return;
            const int rc = kind == Kind::SYRK ? bhblas_syrk(dt, am, ak, ptr(A), ak, ptr(C), am, ctx->stream)
                                              : bhblas_herk(dt, am, ak, ptr(A), ak, ptr(C), am, ctx->stream);
            if (rc != 0) throw std::runtime_error(name + ": " + g_err);
            return;
        }
        const bh_view &B = instr->operand[2];
        require_matrix(B, "B", name);
        if (B.base->dtype() != A.base->dtype()) throw std::runtime_error(name + ": operands differ in type");
        const int64_t bm = B.shape[0], bn = B.shape[1];
        int rc = 0;
        switch (kind) {
            case Kind::GEMM:
                if (bm != ak || C.shape[0] != am || C.shape[1] != bn)
                    shape_error("shapes do not chain (A m x k, B k x n, C m x n)");
                if (ctx->dry_run) return;
                rc = bhblas_gemm(dt, 0, am, bn, ak, ptr(A), ak, ptr(B), bn, ptr(C), bn, ctx->stream);
                break;
            case Kind::GEMMT:
                if (bm != am || C.shape[0] != ak || C.shape[1] != bn)
                    shape_error("shapes do not chain (A m x k, B m x n, C k x n)");
                if (ctx->dry_run) return;
                rc = bhblas_gemm(dt, 1, ak, bn, am, ptr(A), ak, ptr(B), bn, ptr(C), bn, ctx->stream);
                break;
            case Kind::SYMM:
            case Kind::HEMM:
                if (ak != am || bm != am || C.shape[0] != am || C.shape[1] != bn)
                    shape_error("shapes do not chain (A m x m, B m x n, C m x n)");
                if (ctx->dry_run) return;
                rc = (kind == Kind::SYMM ? bhblas_symm : bhblas_hemm)(dt, am, bn, ptr(A), ak, ptr(B), bn, ptr(C), bn,
                                                                      ctx->stream);
                break;
            case Kind::SYR2K:
            case Kind::HER2K:
                if (bm != am || bn != ak || C.shape[0] != am || C.shape[1] != am)
                    shape_error("shapes do not chain (A n x k, B n x k, C n x n)");
                if (ctx->dry_run) return;
                rc = (kind == Kind::SYR2K ? bhblas_syr2k : bhblas_her2k)(dt, am, ak, ptr(A), ak, ptr(B), bn, ptr(C),
                                                                         am, ctx->stream);
                break;
            case Kind::TRMM:
            case Kind::TRSM:
                if (ak != am || bm != am) shape_error("shapes do not chain (A m x m, B m x n)");
                if (ctx->dry_run) return;
                rc = (kind == Kind::TRMM ? bhblas_trmm : bhblas_trsm)(dt, am, bn, ptr(A), ak, ptr(B), bn, ctx->stream);
                break;
            default: break;
        }
        if (rc != 0) throw std::runtime_error(name + ": " + g_err);
    }
};

} // namespace

// Staging area of bhblas_symm_noalloc: the expanded m x m matrix lives here instead of in a
// temporary.  Every device that runs the library holds its own copy (at its own address) until
// the process ends.  In host code `scratch` names a host-side stand-in, never the device memory:
// host code uses only the address that hipGetSymbolAddress returned for the current device.
static constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_address(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        g_err = "bhblas: no usable current device for the staging area";
        return 1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            g_err = std::string("bhblas: hipGetSymbolAddress of the staging area failed: ") + hipGetErrorString(e);
            return 1;
        }
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return 0;
}

} // namespace

extern "C" {

int bhblas_gemm(int dtype, int trans_a, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, const void *B,
                int64_t ldb, void *C, int64_t ldc, void *stream) {
    const Gemm g = make_gemm(dtype, m, n, k, A, lda, B, ldb, C, ldc, stream);
    return trans_a ? dispatch<true, false, false>(g) : dispatch<false, false, false>(g);
}

int bhblas_syrk(int dtype, int64_t m, int64_t k, const void *A, int64_t lda, void *C, int64_t ldc, void *stream) {
    // C = A A^T: op(B)(kk, n) = A[n][kk], i.e. B = A read transposed.
    return dispatch<false, true, true>(make_gemm(dtype, m, m, k, A, lda, A, lda, C, ldc, stream));
}

int bhblas_herk(int dtype, int64_t n, int64_t k, const void *A, int64_t lda, void *C, int64_t ldc, void *stream) {
    if (!is_complex(dtype)) {
        g_err = "bhblas: herk needs a complex type";
        return 1;
    }
    return dispatch<false, true, true, true, true>(make_gemm(dtype, n, n, k, A, lda, A, lda, C, ldc, stream));
}

int bhblas_syr2k(int dtype, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                 int64_t ldc, void *stream) {
    if (dispatch<false, true, true>(make_gemm(dtype, n, n, k, A, lda, B, ldb, C, ldc, stream)) != 0) return 1;
    return dispatch<false, true, true>(make_gemm(dtype, n, n, k, B, ldb, A, lda, C, ldc, stream, 1.0, 1.0));
}

int bhblas_her2k(int dtype, int64_t n, int64_t k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                 int64_t ldc, void *stream) {
    if (!is_complex(dtype)) {
        g_err = "bhblas: her2k needs a complex type";
        return 1;
    }
    if (dispatch<false, true, true, true, false>(make_gemm(dtype, n, n, k, A, lda, B, ldb, C, ldc, stream)) != 0)
        return 1;
    return dispatch<false, true, true, true, true>(make_gemm(dtype, n, n, k, B, ldb, A, lda, C, ldc, stream, 1.0, 1.0));
}

int bhblas_symm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                int64_t ldc, void *stream) {
    return expanded_product<Fill::SYM>(dtype, m, n, A, lda, B, ldb, C, ldc, stream);
}

int bhblas_symm_noalloc(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb,
                        void *C, int64_t ldc, void *stream) {
    if (!check_call(dtype, m, n, m)) return 1;
    if (m == 0 || n == 0) return 0;
    const uint64_t need = static_cast<uint64_t>(m) * static_cast<uint64_t>(m) * static_cast<uint64_t>(elem_size(dtype));
    if (need > kScratchBytes) {
        g_err = "bhblas: symm_noalloc needs m * m * element size <= " + std::to_string(kScratchBytes) +
                " bytes of staging, got " + std::to_string(need) + " (m = " + std::to_string(m) + ")";
        return 1;
    }
    void *staging = nullptr;
    if (scratch_address(&staging) != 0) return 1;
    return expanded_product<Fill::SYM>(dtype, m, n, A, lda, B, ldb, C, ldc, stream, staging);
}

int bhblas_hemm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                int64_t ldc, void *stream) {
    if (!is_complex(dtype)) {
        g_err = "bhblas: hemm needs a complex type";
        return 1;
    }
    return expanded_product<Fill::HERM>(dtype, m, n, A, lda, B, ldb, C, ldc, stream);
}

int bhblas_trmm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *B, int64_t ldb, void *stream) {
    if (!check_call(dtype, m, n, m)) return 1;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t es = elem_size(dtype);
    DeviceTemp prod(s);
    if (!prod.alloc(static_cast<size_t>(m * n * es))) return 1;
    if (expanded_product<Fill::UNIT>(dtype, m, n, A, lda, B, ldb, prod.p, n, stream) != 0) return 1;
    if (hipMemcpy2DAsync(B, static_cast<size_t>(ldb * es), prod.p, static_cast<size_t>(n * es),
                         static_cast<size_t>(n * es), static_cast<size_t>(m), hipMemcpyDeviceToDevice, s) != hipSuccess) {
        (void)hipGetLastError();
        g_err = "bhblas: copy of the trmm product failed";
        return 1;
    }
    return 0;
}

int bhblas_trsm(int dtype, int64_t m, int64_t n, const void *A, int64_t lda, void *B, int64_t ldb, void *stream) {
    if (!check_call(dtype, m, n, m)) return 1;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t es = elem_size(dtype);
    // Diagonal blocks of NB rows from the bottom up: solve the block in place, then take its
    // contribution out of every row above it with one GEMM (C = B above, alpha -1, beta 1).
    const int64_t NB = dtype == BHBLAS_C128 ? 32 : 64;
    const char *a = static_cast<const char *>(A);
    char *b = static_cast<char *>(B);
    for (int64_t r0 = ((m - 1) / NB) * NB; r0 >= 0; r0 -= NB) {
        const int nb = static_cast<int>(std::min(NB, m - r0));
        const void *ad = a + (r0 * lda + r0) * es;
        void *bd = b + r0 * ldb * es;
        switch (dtype) {
            case BHBLAS_F32: trsm_block<float, 64>(ad, lda, bd, ldb, nb, n, s); break;
            case BHBLAS_F64: trsm_block<double, 64>(ad, lda, bd, ldb, nb, n, s); break;
            case BHBLAS_C64: trsm_block<Cplx<float>, 64>(ad, lda, bd, ldb, nb, n, s); break;
            case BHBLAS_C128: trsm_block<Cplx<double>, 32>(ad, lda, bd, ldb, nb, n, s); break;
        }
        if (launch_status() != 0) return 1;
        if (r0 > 0 &&
            dispatch<false, false, false>(make_gemm(dtype, r0, n, nb, a + r0 * es, lda, bd, ldb, B, ldb, stream, -1.0,
                                                    1.0)) != 0)
            return 1;
    }
    return 0;
}

const char *bhblas_last_error(void) { return g_err.c_str(); }

#define BH_BLAS_METHOD(NAME, KIND)                                                                          \
    bh_extmethod_impl *blas_##NAME##_create(void) { return new BlasMethod(Kind::KIND, "blas_" #NAME); }    \
    void blas_##NAME##_destroy(bh_extmethod_impl *self) { delete self; }

BH_BLAS_METHOD(gemm, GEMM)
BH_BLAS_METHOD(gemmt, GEMMT)
BH_BLAS_METHOD(symm, SYMM)
BH_BLAS_METHOD(hemm, HEMM)
BH_BLAS_METHOD(syrk, SYRK)
BH_BLAS_METHOD(herk, HERK)
BH_BLAS_METHOD(syr2k, SYR2K)
BH_BLAS_METHOD(her2k, HER2K)
BH_BLAS_METHOD(trmm, TRMM)
BH_BLAS_METHOD(trsm, TRSM)

} // extern "C"
