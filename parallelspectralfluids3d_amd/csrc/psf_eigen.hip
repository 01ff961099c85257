// Device eigensolver for the basis of obstacle scenes (C ABI in psf.h).
//
// The reference takes its basis from the K smallest accepted eigenpairs of
// the pruned DEC 1-form Laplacian (psfsolver.cpp:267-341: Spectra
// shift-invert, one sparse LU per shift, eigenvalues above 1e-10 kept).
// Here the K smallest eigenpairs of a symmetric sparse matrix A (the
// caller's CSR) come from Chebyshev-filtered subspace iteration on a block
// of m = K + guard vectors, all on the GPU:
//   start  X = orth(R), R Gaussian;
//   repeat Rayleigh-Ritz (H = X^T A X, H = V diag(w) V^T, X <- X V), stop
//          when the K smallest residuals ||A x - w x|| <= tol * lambda_max;
//          else X <- orth(p(A) X), p the Chebyshev polynomial of the given
//          degree that is at most 1 on [w_m, lambda_max] and grows below it.
// A null space is not special: its vectors are the smallest eigenpairs.  The
// caller lifts one it does not want (the DEC Laplacian's gradients) with a
// penalty term, as the host shift-invert path does.
// Products with A: one fused CSR kernel per polynomial degree (k_csr_block2
// with 16-byte column-pair loads for an even m, else k_csr_block).  Gram matrices:
// split-reduction DGEMM (psf_blas.hpp); rotations: rocBLAS DGEMM; orthonormalisation: Cholesky QR twice
// (rocSOLVER POTRF, rocBLAS TRSM in place); H's eigendecomposition: rocSOLVER
// SYEVD.  The block is stored row-major (n x m), so each nonzero of a CSR row
// reads one contiguous row of the block; for the BLAS calls that is the
// column-major m x n matrix X^T.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "psf_blas.hpp"
#include "psf_internal.hpp"

namespace psf {
namespace {

constexpr uint32_t kMaxBlock = 1024;

int set_err(psf_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return ctx_set_error(ctx, code, buf);
}

#define EIG_HIPCHK(ctx, expr)                                                                                 \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return set_err((ctx), PSF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)
#define EIG_BLASCHK(ctx, expr)                                                                                \
    do {                                                                                                      \
        rocblas_status s_ = (expr);                                                                           \
        if (s_ != rocblas_status_success)                                                                     \
            return set_err((ctx), PSF_ERR_HIP, "%s: rocBLAS status %d", #expr, static_cast<int>(s_));         \
    } while (0)

// Launch blocks go to the 8 XCDs round-robin (block b to XCD b % 8), each
// with its own L2: logical block (b % 8) * per + b / 8 gives every XCD a
// contiguous range of rows, whose neighbours (rows in spatial order) then
// stay in that XCD's L2.
constexpr uint32_t kXcds = 8;

// out = alpha * (A Y) + beta * Y + gamma * Z (Z may be null) on the m columns
// starting at each pointer, rows ld apart: one wave per row; lane l holds
// columns l, l + 64, ..., l + 64 (R - 1) (R = ceil(m / 64)), so each nonzero
// issues R independent loads of one block row.  (Running the Chebyshev
// recurrence on one 64-column slice at a time, R = 1, was 30% slower per
// product than the whole block at R = 7 on the 64^3 obstacle scene.)
template <int R>
__global__ __launch_bounds__(256) void k_csr_block(uint32_t n, uint32_t m, uint32_t ld, uint32_t per,
                                                   const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                   const double* __restrict__ av, const double* __restrict__ Y,
                                                   const double* __restrict__ Z, double alpha, double beta,
                                                   double gamma, double* __restrict__ out)
{
    const uint32_t blk = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    const uint32_t row = __builtin_amdgcn_readfirstlane(blk * 4u + (threadIdx.x >> 6));
    if (row >= n) return;
    const uint32_t lane = threadIdx.x & 63u;
    const int32_t b = rp[row], e = rp[row + 1];
    double s[R];
#pragma unroll
    for (int r = 0; r < R; ++r) s[r] = 0.0;
    // four nonzeros' loads are issued before their products are summed (in
    // the same order as one at a time)
    int32_t q = b;
    for (; q + 4 <= e; q += 4) {
        double a[4];
        const double* y[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[j] = av[q + j];
            y[j] = Y + static_cast<size_t>(ci[q + j]) * ld + lane;
        }
        double v[4][R];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < R; ++r) v[j][r] = lane + 64u * r < m ? y[j][64 * r] : 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (lane + 64u * r < m) s[r] += a[j] * v[j][r];
    }
    for (; q < e; ++q) {
        const double a = av[q];
        const double* y = Y + static_cast<size_t>(ci[q]) * ld + lane;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lane + 64u * r < m) s[r] += a * y[64 * r];
    }
    const size_t o = static_cast<size_t>(row) * ld + lane;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (lane + 64u * r >= m) break;
        double v = alpha * s[r] + beta * Y[o + 64 * r];
        if (Z) v += gamma * Z[o + 64 * r];
        out[o + 64 * r] = v;
    }
}

template <int R>
void launch_csr(hipStream_t st, uint32_t n, uint32_t m, uint32_t ld, const int32_t* rp, const int32_t* ci,
                const double* av, const double* Y, const double* Z, double a, double b, double g, double* out)
{
    const uint32_t per = ((n + 3) / 4 + kXcds - 1) / kXcds;
    k_csr_block<R><<<per * kXcds, 256, 0, st>>>(n, m, ld, per, rp, ci, av, Y, Z, a, b, g, out);
}

using CsrLaunch = void (*)(hipStream_t, uint32_t, uint32_t, uint32_t, const int32_t*, const int32_t*, const double*,
                           const double*, const double*, double, double, double, double*);
constexpr CsrLaunch kCsrLaunch[16] = {launch_csr<1>,  launch_csr<2>,  launch_csr<3>,  launch_csr<4>,
                                      launch_csr<5>,  launch_csr<6>,  launch_csr<7>,  launch_csr<8>,
                                      launch_csr<9>,  launch_csr<10>, launch_csr<11>, launch_csr<12>,
                                      launch_csr<13>, launch_csr<14>, launch_csr<15>, launch_csr<16>};

// The same for an even m, with rows ld = m apart: lane l holds the column
// pairs (2l, 2l + 1) + 128 r, r < R2 = ceil(m / 128), each loaded as one
// 16-byte load, so a nonzero costs R2 load instructions instead of
// ceil(m / 64) (even m keeps every row start and pair 16-byte aligned)
template <int R2>
__global__ __launch_bounds__(256) void k_csr_block2(uint32_t n, uint32_t m, uint32_t per,
                                                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                    const double* __restrict__ av, const double* __restrict__ Y,
                                                    const double* __restrict__ Z, double alpha, double beta,
                                                    double gamma, double* __restrict__ out)
{
    const uint32_t blk = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    const uint32_t row = __builtin_amdgcn_readfirstlane(blk * 4u + (threadIdx.x >> 6));
    if (row >= n) return;
    const uint32_t c0 = 2u * (threadIdx.x & 63u);
    const int32_t b = rp[row], e = rp[row + 1];
    double2 s[R2];
#pragma unroll
    for (int r = 0; r < R2; ++r) s[r] = make_double2(0.0, 0.0);
    auto add = [&](double a, const double2& v, double2& acc) {
        acc.x += a * v.x;
        acc.y += a * v.y;
    };
    int32_t q = b;
    for (; q + 4 <= e; q += 4) {
        double a[4];
        const double2* y[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[j] = av[q + j];
            y[j] = reinterpret_cast<const double2*>(Y + static_cast<size_t>(ci[q + j]) * m + c0);
        }
        double2 v[4][R2];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < R2; ++r) v[j][r] = c0 + 128u * r < m ? y[j][64 * r] : make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < R2; ++r)
                if (c0 + 128u * r < m) add(a[j], v[j][r], s[r]);
    }
    for (; q < e; ++q) {
        const double a = av[q];
        const double2* y = reinterpret_cast<const double2*>(Y + static_cast<size_t>(ci[q]) * m + c0);
#pragma unroll
        for (int r = 0; r < R2; ++r)
            if (c0 + 128u * r < m) add(a, y[64 * r], s[r]);
    }
    const size_t o = static_cast<size_t>(row) * m + c0;
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        if (c0 + 128u * r >= m) break;
        const double2 yv = *reinterpret_cast<const double2*>(Y + o + 128 * r);
        double2 v;
        v.x = alpha * s[r].x + beta * yv.x;
        v.y = alpha * s[r].y + beta * yv.y;
        if (Z) {
            const double2 zv = *reinterpret_cast<const double2*>(Z + o + 128 * r);
            v.x += gamma * zv.x;
            v.y += gamma * zv.y;
        }
        *reinterpret_cast<double2*>(out + o + 128 * r) = v;
    }
}

template <int R2>
void launch_csr2(hipStream_t st, uint32_t n, uint32_t m, uint32_t, const int32_t* rp, const int32_t* ci,
                 const double* av, const double* Y, const double* Z, double a, double b, double g, double* out)
{
    const uint32_t per = ((n + 3) / 4 + kXcds - 1) / kXcds;
    k_csr_block2<R2><<<per * kXcds, 256, 0, st>>>(n, m, per, rp, ci, av, Y, Z, a, b, g, out);
}

constexpr CsrLaunch kCsr2Launch[8] = {launch_csr2<1>, launch_csr2<2>, launch_csr2<3>, launch_csr2<4>,
                                      launch_csr2<5>, launch_csr2<6>, launch_csr2<7>, launch_csr2<8>};

// Gaussian start block: Box-Muller on two SplitMix64 draws of (seed, index)
__device__ __forceinline__ uint64_t splitmix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gaussian(uint64_t seed, size_t count, double* __restrict__ out)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= count) return;
    const uint64_t a = splitmix(seed ^ (2 * t)), b = splitmix(seed ^ (2 * t + 1));
    const double u1 = (static_cast<double>(a >> 11) + 1.0) * 0x1p-53;   // (0, 1]
    const double u2 = static_cast<double>(b >> 11) * 0x1p-53;            // [0, 1)
    out[t] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// part[b * k + c] = sum over rows r = b, b + grid, ... of
// (AX[r][c] - w[c] X[r][c])^2, for c < k
__global__ __launch_bounds__(256) void k_residual_part(uint32_t n, uint32_t m, uint32_t k, const double* __restrict__ AX,
                                                       const double* __restrict__ X, const double* __restrict__ w,
                                                       double* __restrict__ part)
{
    for (uint32_t c = threadIdx.x; c < k; c += 256u) {
        const double wc = w[c];
        double s = 0.0;
        for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
            const double d = AX[static_cast<size_t>(r) * m + c] - wc * X[static_cast<size_t>(r) * m + c];
            s += d * d;
        }
        part[static_cast<size_t>(blockIdx.x) * k + c] = s;
    }
}

__global__ __launch_bounds__(256) void k_residual_sum(uint32_t k, uint32_t parts, const double* __restrict__ part,
                                                      double* __restrict__ sq)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (uint32_t b = 0; b < parts; ++b) s += part[static_cast<size_t>(b) * k + c];
    sq[c] = s;
}

// out[j n + i] = X[i m + j] for i < n, j < k: the first k columns of the
// row-major n x m block as a column-major n x k matrix, through 32 x 32 LDS
// tiles so that both sides are read and written along rows
__global__ __launch_bounds__(256) void k_take_columns(uint32_t n, uint32_t m, uint32_t k, const double* __restrict__ X,
                                                      double* __restrict__ out)
{
    __shared__ double tile[32][33];
    const uint32_t i0 = blockIdx.x * 32u, j0 = blockIdx.y * 32u;
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;   // ty in 0..7
    for (uint32_t r = ty; r < 32u; r += 8u) {
        const uint32_t i = i0 + r, j = j0 + tx;
        if (i < n && j < k) tile[r][tx] = X[static_cast<size_t>(i) * m + j];
    }
    __syncthreads();
    for (uint32_t r = ty; r < 32u; r += 8u) {
        const uint32_t j = j0 + r, i = i0 + tx;
        if (i < n && j < k) out[static_cast<size_t>(j) * n + i] = tile[tx][r];
    }
}

// H <- (H + H^T) / 2 for the column-major m x m H
__global__ void k_symmetrise(uint32_t m, double* __restrict__ H)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < m && i > j) {
        const double v = 0.5 * (H[static_cast<size_t>(j) * m + i] + H[static_cast<size_t>(i) * m + j]);
        H[static_cast<size_t>(j) * m + i] = v;
        H[static_cast<size_t>(i) * m + j] = v;
    }
}

struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t n)
    {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

struct Blas {
    rocblas_handle h = nullptr;
    ~Blas()
    {
        if (h) (void)rocblas_destroy_handle(h);
    }
};

}  // namespace
}  // namespace psf

using namespace psf;

extern "C" {

int psf_eigensolve_csr(psf_ctx* ctx, uint32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* values,
                       uint32_t K, uint32_t guard, uint32_t degree, double tol, uint32_t max_products, uint64_t seed,
                       double* lambda_out, double* X_out, uint32_t* products_out)
{
    if (!ctx) return PSF_ERR_ARG;
    if (!row_ptr || !col_idx || !values || !lambda_out || !X_out) return set_err(ctx, PSF_ERR_ARG, "null argument");
    const uint64_t m64 = uint64_t(K) + guard;
    if (K == 0 || guard == 0 || m64 > kMaxBlock || m64 >= n || degree == 0 || degree > 1000 || !(tol > 0.0) ||
        n > (1u << 30))
        return set_err(ctx, PSF_ERR_ARG,
                       "eigensolve: need K >= 1, guard >= 1, K + guard <= %u and < n, 1 <= degree <= 1000, tol > 0",
                       kMaxBlock);
    const uint32_t m = static_cast<uint32_t>(m64);
    if (row_ptr[0] != 0) return set_err(ctx, PSF_ERR_ARG, "eigensolve: row_ptr[0] must be 0");
    const int32_t nnz = row_ptr[n];
    if (nnz <= 0) return set_err(ctx, PSF_ERR_ARG, "eigensolve: empty matrix");
    // Gershgorin bound of the top of the spectrum, and the CSR checks
    double lmax = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) return set_err(ctx, PSF_ERR_ARG, "eigensolve: row pointers decrease at %u", i);
        double s = 0.0;
        for (int32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
            if (col_idx[q] < 0 || static_cast<uint32_t>(col_idx[q]) >= n)
                return set_err(ctx, PSF_ERR_ARG, "eigensolve: column index out of range in row %u", i);
            if (!std::isfinite(values[q])) return set_err(ctx, PSF_ERR_ARG, "eigensolve: non-finite value in row %u", i);
            s += std::fabs(values[q]);
        }
        lmax = std::max(lmax, s);
    }
    if (!(lmax > 0.0)) return set_err(ctx, PSF_ERR_ARG, "eigensolve: zero matrix");
    EIG_HIPCHK(ctx, hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    const size_t blk = size_t(n) * m;
    const uint32_t rparts = std::min<uint32_t>(1024, n);
    DevMem mem;
    int32_t *drp, *dci, *dinfo;
    double *dav, *X, *Y, *P, *Q, *H, *dw, *dE, *dpart, *dsq, *dsplit;
    EIG_HIPCHK(ctx, mem.alloc(&drp, size_t(n) + 1));
    EIG_HIPCHK(ctx, mem.alloc(&dci, size_t(nnz)));
    EIG_HIPCHK(ctx, mem.alloc(&dav, size_t(nnz)));
    EIG_HIPCHK(ctx, mem.alloc(&X, blk));
    EIG_HIPCHK(ctx, mem.alloc(&Y, blk));
    EIG_HIPCHK(ctx, mem.alloc(&P, blk));
    EIG_HIPCHK(ctx, mem.alloc(&Q, blk));
    EIG_HIPCHK(ctx, mem.alloc(&H, size_t(m) * m));
    EIG_HIPCHK(ctx, mem.alloc(&dw, m));
    EIG_HIPCHK(ctx, mem.alloc(&dE, m));
    EIG_HIPCHK(ctx, mem.alloc(&dpart, size_t(rparts) * K));
    EIG_HIPCHK(ctx, mem.alloc(&dsq, K));
    EIG_HIPCHK(ctx, mem.alloc(&dinfo, 1));
    EIG_HIPCHK(ctx, mem.alloc(&dsplit, splitk_work_doubles(static_cast<int>(m), static_cast<int>(m), n)));
    EIG_HIPCHK(ctx, hipMemcpy(drp, row_ptr, (size_t(n) + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    EIG_HIPCHK(ctx, hipMemcpy(dci, col_idx, size_t(nnz) * sizeof(int32_t), hipMemcpyHostToDevice));
    EIG_HIPCHK(ctx, hipMemcpy(dav, values, size_t(nnz) * sizeof(double), hipMemcpyHostToDevice));
    k_gaussian<<<static_cast<unsigned>((blk + 255) / 256), 256, 0, st>>>(seed, blk, Y);
    EIG_HIPCHK(ctx, hipGetLastError());
    Blas blas;
    EIG_BLASCHK(ctx, rocblas_create_handle(&blas.h));
    EIG_BLASCHK(ctx, rocblas_set_stream(blas.h, st));
    const rocblas_int mi = static_cast<rocblas_int>(m), ni = static_cast<rocblas_int>(n);
    const double one = 1.0, zero = 0.0;
    uint32_t products = 0;
    const CsrLaunch csr = m % 2 == 0 ? kCsr2Launch[(m + 127) / 128 - 1] : kCsrLaunch[(m + 63) / 64 - 1];
    auto apply = [&](const double* in, const double* z, double a, double b, double g, double* out) {
        csr(st, n, m, m, drp, dci, dav, in, z, a, b, g, out);
        ++products;
        return hipGetLastError();
    };
    int32_t info = 0;
    // x <- x R^-1 with x^T x = R^T R, twice (Cholesky QR2)
    auto orthonormalise = [&](double* x) -> int {
        for (int pass = 0; pass < 2; ++pass) {
            EIG_BLASCHK(ctx, gemm_splitk(blas.h, rocblas_operation_none, rocblas_operation_transpose, mi, mi, ni, x, mi,
                                         mi, x, mi, mi, H, mi, dsplit));
            EIG_BLASCHK(ctx, rocsolver_dpotrf(blas.h, rocblas_fill_upper, mi, H, mi, dinfo));
            EIG_HIPCHK(ctx, hipMemcpyAsync(&info, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            EIG_HIPCHK(ctx, hipStreamSynchronize(st));
            if (info != 0)
                return set_err(ctx, PSF_ERR_NONFINITE, "eigensolve: the block lost rank (Cholesky pivot %d)", info);
            // (x R^-1)^T = R^-T x^T: a left solve with R^T on the column-major x^T
            EIG_BLASCHK(ctx, rocblas_dtrsm(blas.h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                           rocblas_diagonal_non_unit, mi, ni, &one, H, mi, x, mi));
        }
        return PSF_OK;
    };
    // out = in V (row-major n x m times the column-major m x m V in H):
    // out^T = V^T in^T
    auto rotate = [&](const double* in, double* out) -> int {
        EIG_BLASCHK(ctx, rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, mi, ni, mi, &one, H,
                                       mi, in, mi, &zero, out, mi));
        return PSF_OK;
    };
    // X, Y, P and Q always name the four block buffers in some order
    std::swap(X, Y);   // the Gaussian start
    if (int rc = orthonormalise(X)) return rc;
    std::vector<double> w(m), sq(K);
    bool converged = false;
    double worst = 0.0;
    for (;;) {
        // Rayleigh-Ritz: Y = A X, H = X^T Y, then X <- X V and Y <- Y V
        EIG_HIPCHK(ctx, apply(X, nullptr, 1.0, 0.0, 0.0, Y));
        EIG_BLASCHK(ctx, gemm_splitk(blas.h, rocblas_operation_none, rocblas_operation_transpose, mi, mi, ni, X, mi, mi,
                                     Y, mi, mi, H, mi, dsplit));
        k_symmetrise<<<dim3((m + 255) / 256, m), 256, 0, st>>>(m, H);
        EIG_HIPCHK(ctx, hipGetLastError());
        EIG_BLASCHK(ctx, rocsolver_dsyevd(blas.h, rocblas_evect_original, rocblas_fill_upper, mi, H, mi, dw, dE, dinfo));
        EIG_HIPCHK(ctx, hipMemcpyAsync(&info, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        EIG_HIPCHK(ctx, hipMemcpyAsync(w.data(), dw, m * sizeof(double), hipMemcpyDeviceToHost, st));
        if (int rc = rotate(X, P)) return rc;
        std::swap(X, P);
        if (int rc = rotate(Y, P)) return rc;
        std::swap(Y, P);
        k_residual_part<<<rparts, 256, 0, st>>>(n, m, K, Y, X, dw, dpart);
        EIG_HIPCHK(ctx, hipGetLastError());
        k_residual_sum<<<(K + 255) / 256, 256, 0, st>>>(K, rparts, dpart, dsq);
        EIG_HIPCHK(ctx, hipGetLastError());
        EIG_HIPCHK(ctx, hipMemcpyAsync(sq.data(), dsq, K * sizeof(double), hipMemcpyDeviceToHost, st));
        EIG_HIPCHK(ctx, hipStreamSynchronize(st));
        if (info != 0) return set_err(ctx, PSF_ERR_NONFINITE, "eigensolve: SYEVD did not converge (%d)", info);
        worst = 0.0;
        for (uint32_t j = 0; j < K; ++j) worst = std::max(worst, std::sqrt(sq[j]));
        if (!std::isfinite(worst)) return set_err(ctx, PSF_ERR_NONFINITE, "eigensolve: non-finite residual");
        if (worst <= tol * lmax) {
            converged = true;
            break;
        }
        if (uint64_t(products) + degree + 2 > max_products) break;
        // Chebyshev filter of the given degree on [w[m-1], lmax]:
        // T1 = (A X - c X) / e, T(k+1) = 2 (A Tk - c Tk) / e - T(k-1)
        const double lo = w[m - 1], c = 0.5 * (lmax + lo), e = 0.5 * (lmax - lo);
        if (!(e > 0.0)) return set_err(ctx, PSF_ERR_NONFINITE, "eigensolve: empty filter interval");
        double* t[3] = {X, P, Q};   // T(k-1), T(k), next; rotates through X, P, Q
        EIG_HIPCHK(ctx, apply(t[0], nullptr, 1.0 / e, -c / e, 0.0, t[1]));
        for (uint32_t d = 2; d <= degree; ++d) {
            EIG_HIPCHK(ctx, apply(t[1], t[0], 2.0 / e, -2.0 * c / e, -1.0, t[2]));
            std::rotate(t, t + 1, t + 3);
        }
        if (t[1] != X) std::swap(X, t[1] == P ? P : Q);   // X <- T(degree); the old X is scratch
        if (int rc = orthonormalise(X)) return rc;
    }
    if (products_out) *products_out = products;
    if (!converged)
        return set_err(ctx, PSF_ERR_STATE, "eigensolve: residual %.3e x lambda_max after %u products", worst / lmax,
                       products);
    // eigenvalues and the K smallest vectors, n x K column-major (P is free
    // scratch of n x m >= n x K doubles)
    k_take_columns<<<dim3((n + 31) / 32, (K + 31) / 32), 256, 0, st>>>(n, m, K, X, P);
    EIG_HIPCHK(ctx, hipGetLastError());
    EIG_HIPCHK(ctx, hipMemcpyAsync(X_out, P, size_t(n) * K * sizeof(double), hipMemcpyDeviceToHost, st));
    EIG_HIPCHK(ctx, hipStreamSynchronize(st));
    for (uint32_t j = 0; j < K; ++j) lambda_out[j] = w[j];
    return PSF_OK;
}

}  // extern "C"
