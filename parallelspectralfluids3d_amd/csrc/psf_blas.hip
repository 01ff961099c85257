// Split-reduction DGEMM (psf_blas.hpp).
#include "psf_blas.hpp"

#include <algorithm>

namespace psf {
namespace {

constexpr int kTileMN = 128;          // output tile of rocBLAS's FP64 kernels
constexpr int kTargetGroups = 1024;   // workgroups to aim for (4 per CU)
constexpr long long kMinChunk = 512;  // shortest reduction chunk worth a partial

int split_count(int m, int n, long long k)
{
    const long long tiles = static_cast<long long>((m + kTileMN - 1) / kTileMN) * ((n + kTileMN - 1) / kTileMN);
    long long s = (kTargetGroups + tiles - 1) / tiles;
    s = std::min(s, std::max(1LL, k / kMinChunk));
    return static_cast<int>(std::max(1LL, std::min(s, 1024LL)));
}

// C[i + j ldc] = sum over s < S of P[s m n + i + j m]
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ P, int S, int m, int n,
                                                      double* __restrict__ C, int ldc)
{
    const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const long long mn = static_cast<long long>(m) * n;
    if (t >= mn) return;
    double s = 0.0;
    for (int q = 0; q < S; ++q) s += P[q * mn + t];
    const long long i = t % m, j = t / m;
    C[i + j * ldc] = s;
}

}  // namespace

size_t splitk_work_doubles(int m, int n, long long k)
{
    return static_cast<size_t>(split_count(m, n, k)) * m * n;
}

rocblas_status gemm_splitk(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, long long k,
                           const double* A, int lda, long long a_step, const double* B, int ldb, long long b_step,
                           double* C, int ldc, double* work)
{
    const double one = 1.0, zero = 0.0;
    const int S = split_count(m, n, k);
    if (S == 1)
        return rocblas_dgemm(h, ta, tb, m, n, static_cast<rocblas_int>(k), &one, A, lda, B, ldb, &zero, C, ldc);
    const long long kc = k / S;            // chunks 0 .. S-2 have kc; the last takes the rest
    const long long mn = static_cast<long long>(m) * n;
    rocblas_status st = rocblas_dgemm_strided_batched(h, ta, tb, m, n, static_cast<rocblas_int>(kc), &one, A, lda,
                                                      kc * a_step, B, ldb, kc * b_step, &zero, work, m, mn, S - 1);
    if (st != rocblas_status_success) return st;
    const long long k0 = (S - 1) * kc;
    st = rocblas_dgemm(h, ta, tb, m, n, static_cast<rocblas_int>(k - k0), &one, A + k0 * a_step, lda, B + k0 * b_step,
                       ldb, &zero, work + (S - 1) * mn, m);
    if (st != rocblas_status_success) return st;
    hipStream_t s = nullptr;
    st = rocblas_get_stream(h, &s);
    if (st != rocblas_status_success) return st;
    k_sum_partials<<<static_cast<unsigned>((mn + 255) / 256), 256, 0, s>>>(work, S, m, n, C, ldc);
    return hipGetLastError() == hipSuccess ? rocblas_status_success : rocblas_status_internal_error;
}

}  // namespace psf
