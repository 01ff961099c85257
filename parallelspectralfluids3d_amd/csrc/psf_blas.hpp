// Split-reduction DGEMM for products with a small output and a long
// reduction (Gram matrices of tall blocks, C_i = V^T W_i): a plain rocBLAS
// call tiles only the m x n output and runs a handful of workgroups, so the
// reduction is cut into S chunks, computed as one strided-batched DGEMM into
// S partial m x n products, and summed by a small kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

namespace psf {

// C (column-major m x n, leading dimension ldc) = op(A) op(B), reduction
// length k.  a_step / b_step: elements by which A / B advance per unit of the
// reduction index (1 for a transposed column-major operand, its leading
// dimension for a plain one).  work: at least splitk_work_doubles(m, n, k)
// doubles.  Runs on the handle's stream.
size_t splitk_work_doubles(int m, int n, long long k);
rocblas_status gemm_splitk(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, long long k,
                           const double* A, int lda, long long a_step, const double* B, int ldb, long long b_step,
                           double* C, int ldc, double* work);

}  // namespace psf
