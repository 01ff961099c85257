// Structure tensor C from any basis, built on the device (C ABI in psf.h).
//
// PSFSolver::buildAdvection (psfsolver.cpp:431-551) sums, for every inside
// voxel and each of its 12 edge slots s (edge e, faces a and b),
//     W_i[e][j] += s_h c(a, b) (Phi[a][i] Phi[b][j] - Phi[a][j] Phi[b][i])
// into K wedge matrices W_i (E x K), then C_i = (1/lambda_i) V^T W_i with
// V = curl Phi (E x K).  The caller hands over the per-(voxel, slot) terms
// (edge, a, b, s_h c(a, b)) and V as the reference computes them; here:
//   - the terms are grouped by edge on the host (a CSR over edges);
//   - k_wedge_rows writes W_i / lambda_i row by row for a chunk of modes i,
//     one wave per (edge, i), the K columns j across the lanes, four per
//     lane (rows of Phi are read contiguously from a row-major copy);
//   - a split-reduction DGEMM (psf_blas.hpp) per mode gives
//     C_i = V^T (W_i / lambda_i) on the FP64 matrix cores;
//   - C (FP64) is then stored in the scene's precision like an upload.
// Memory: Phi and V once, plus one chunk of wedge matrices sized to a few
// GB, instead of the reference's 2 K E K doubles.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "psf_blas.hpp"
#include "psf_internal.hpp"
#include "psf_kernels.hpp"
#include "psf_persistent.hpp"

namespace psf {
namespace {

constexpr size_t kChunkBytes = size_t(4) << 30;   // wedge matrices held at once

int set_err(psf_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return ctx_set_error(ctx, code, buf);
}

#define ADV_HIPCHK(ctx, expr)                                                                                 \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return set_err((ctx), e_ == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP, "%s: %s", #expr,   \
                           hipGetErrorString(e_));                                                            \
    } while (0)
#define ADV_BLASCHK(ctx, expr)                                                                                \
    do {                                                                                                      \
        rocblas_status s_ = (expr);                                                                           \
        if (s_ != rocblas_status_success)                                                                     \
            return set_err((ctx), PSF_ERR_HIP, "%s: rocBLAS status %d", #expr, static_cast<int>(s_));         \
    } while (0)

// row-major F x K copy of the column-major Phi
__global__ __launch_bounds__(256) void k_rows_of(const double* __restrict__ cm, size_t F, uint32_t K,
                                                  double* __restrict__ rm)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= F * K) return;
    const size_t f = t / K, k = t % K;
    rm[t] = cm[k * F + f];
}

// W[(i - i0)][e][j] = (1 / lambda_i) sum over the terms t of edge e of
// coef_t (Phi[a_t][i] Phi[b_t][j] - Phi[a_t][j] Phi[b_t][i]), for the modes
// i0 <= i < i0 + gridDim.y; one wave per (e, i)
__global__ __launch_bounds__(256) void k_wedge_rows(uint32_t E, uint32_t K, uint32_t i0, const uint32_t* __restrict__ ptr,
                                                    const uint32_t* __restrict__ ta, const uint32_t* __restrict__ tb,
                                                    const double* __restrict__ tc, const double* __restrict__ PhiR,
                                                    const double* __restrict__ lambda, double* __restrict__ W)
{
    const uint32_t e = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6));
    if (e >= E) return;
    const uint32_t i = i0 + blockIdx.y;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t t0 = ptr[e], t1 = ptr[e + 1];
    const double inv = 1.0 / lambda[i];
    double* out = W + (static_cast<size_t>(blockIdx.y) * E + e) * K;
    // 256 columns per pass, four per lane: a term's Phi(a, i), Phi(b, i) and
    // coefficient are loaded once for all four, and its eight row loads are
    // in flight together
    for (uint32_t j0 = 0; j0 < K; j0 += 256u) {
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t t = t0; t < t1; ++t) {
            const double* pa = PhiR + static_cast<size_t>(ta[t]) * K;
            const double* pb = PhiR + static_cast<size_t>(tb[t]) * K;
            const double c = tc[t], ai = pa[i], bi = pb[i];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t j = j0 + lane + 64u * r;
                if (j < K) w[r] += c * (ai * pb[j] - pa[j] * bi);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t j = j0 + lane + 64u * r;
            if (j < K) out[j] = w[r] * inv;
        }
    }
}

__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t < n) b[t] = static_cast<float>(a[t]);
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

uint64_t edge_count(const psf_scene_desc& d)
{
    const uint64_t nx = d.nx, ny = d.ny, nz = d.nz;
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
}

}  // namespace
}  // namespace psf

using namespace psf;

extern "C" {

int psf_scene_build_advection(psf_scene* s, const double* Phi, const double* V, uint32_t nE, const double* lambda,
                              uint32_t nterms, const uint32_t* term_edge, const uint32_t* term_a,
                              const uint32_t* term_b, const double* term_coef)
{
    if (!s) return PSF_ERR_ARG;
    psf_ctx* ctx = s->ctx;
    if (!Phi || !V || !lambda || (nterms && (!term_edge || !term_a || !term_b || !term_coef)))
        return set_err(ctx, PSF_ERR_ARG, "null argument");
    const int prec = s->d.c_precision;
    if (prec == PSF_C_FP16) return set_err(ctx, PSF_ERR_UNSUPPORTED, "FP16 C needs per-mode scales: upload it instead");
    const uint32_t K = s->d.K, F = s->F;
    if (uint64_t(nE) != edge_count(s->d))
        return set_err(ctx, PSF_ERR_ARG, "V has %u rows, the grid has %llu edges", nE,
                       static_cast<unsigned long long>(edge_count(s->d)));
    for (uint32_t k = 0; k < K; ++k)
        if (!(lambda[k] != 0.0) || !std::isfinite(lambda[k]))
            return set_err(ctx, PSF_ERR_ARG, "lambda[%u] must be finite and non-zero", k);
    // terms grouped by edge (counting sort, stable)
    std::vector<uint32_t> ptr(size_t(nE) + 1, 0u);
    for (uint32_t t = 0; t < nterms; ++t) {
        if (term_edge[t] >= nE || term_a[t] >= F || term_b[t] >= F)
            return set_err(ctx, PSF_ERR_ARG, "term %u: edge or face index out of range", t);
        if (!std::isfinite(term_coef[t])) return set_err(ctx, PSF_ERR_ARG, "term %u: non-finite coefficient", t);
        ptr[term_edge[t] + 1]++;
    }
    for (uint32_t e = 0; e < nE; ++e) ptr[e + 1] += ptr[e];
    std::vector<uint32_t> ta(nterms), tb(nterms), pos(ptr.begin(), ptr.end() - 1);
    std::vector<double> tc(nterms);
    for (uint32_t t = 0; t < nterms; ++t) {
        const uint32_t p = pos[term_edge[t]]++;
        ta[p] = term_a[t];
        tb[p] = term_b[t];
        tc[p] = term_coef[t];
    }
    ADV_HIPCHK(ctx, hipSetDevice(ctx->device));
    if (int rc = scene_window_flush(s)) return rc;
    // an FP64 C is written in place: until the build completes, the scene has none
    s->has_C = false;
    const hipStream_t st = ctx->stream;
    DevMem mem;
    uint32_t *dptr, *dta, *dtb;
    double *dtc, *dPhi, *dPhiR, *dV, *dlam, *dW, *Cf, *dsplit;
    ADV_HIPCHK(ctx, mem.alloc(&dptr, ptr.size()));
    ADV_HIPCHK(ctx, mem.alloc(&dta, nterms));
    ADV_HIPCHK(ctx, mem.alloc(&dtb, nterms));
    ADV_HIPCHK(ctx, mem.alloc(&dtc, nterms));
    ADV_HIPCHK(ctx, mem.alloc(&dPhi, size_t(F) * K));
    ADV_HIPCHK(ctx, mem.alloc(&dPhiR, size_t(F) * K));
    ADV_HIPCHK(ctx, mem.alloc(&dV, size_t(nE) * K));
    ADV_HIPCHK(ctx, mem.alloc(&dlam, K));
    const size_t row_bytes = size_t(nE) * K * sizeof(double);   // one wedge matrix
    const uint32_t ic = static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>(K, kChunkBytes / row_bytes)));
    ADV_HIPCHK(ctx, mem.alloc(&dW, size_t(ic) * nE * K));
    ADV_HIPCHK(ctx, mem.alloc(&dsplit, splitk_work_doubles(static_cast<int>(K), static_cast<int>(K), nE)));
    const size_t n = size_t(K) * K * K;
    if (prec == PSF_C_FP64) {
        Cf = static_cast<double*>(s->dev.C);
    } else {
        ADV_HIPCHK(ctx, mem.alloc(&Cf, n));
    }
    ADV_HIPCHK(ctx, hipMemcpyAsync(dptr, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice, st));
    ADV_HIPCHK(ctx, hipMemcpyAsync(dta, ta.data(), size_t(nterms) * 4, hipMemcpyHostToDevice, st));
    ADV_HIPCHK(ctx, hipMemcpyAsync(dtb, tb.data(), size_t(nterms) * 4, hipMemcpyHostToDevice, st));
    ADV_HIPCHK(ctx, hipMemcpyAsync(dtc, tc.data(), size_t(nterms) * 8, hipMemcpyHostToDevice, st));
    ADV_HIPCHK(ctx, hipMemcpyAsync(dPhi, Phi, size_t(F) * K * 8, hipMemcpyHostToDevice, st));
    ADV_HIPCHK(ctx, hipMemcpyAsync(dV, V, size_t(nE) * K * 8, hipMemcpyHostToDevice, st));
    ADV_HIPCHK(ctx, hipMemcpyAsync(dlam, lambda, size_t(K) * 8, hipMemcpyHostToDevice, st));
    k_rows_of<<<static_cast<unsigned>((size_t(F) * K + 255) / 256), 256, 0, st>>>(dPhi, F, K, dPhiR);
    ADV_HIPCHK(ctx, hipGetLastError());
    Blas blas;
    ADV_BLASCHK(ctx, rocblas_create_handle(&blas.h));
    ADV_BLASCHK(ctx, rocblas_set_stream(blas.h, st));
    const rocblas_int Ki = static_cast<rocblas_int>(K), Ei = static_cast<rocblas_int>(nE);
    for (uint32_t i0 = 0; i0 < K; i0 += ic) {
        const uint32_t cnt = std::min(ic, K - i0);
        k_wedge_rows<<<dim3((nE + 3) / 4, cnt), 256, 0, st>>>(nE, K, i0, dptr, dta, dtb, dtc, dPhiR, dlam, dW);
        ADV_HIPCHK(ctx, hipGetLastError());
        // C_i (column-major K x K: element (k, j) = C_i(k, j)) = V^T W_i, with
        // V column-major E x K and W_i row-major E x K = column-major K x E;
        // the output is small and the reduction long, so it is split
        for (uint32_t q = 0; q < cnt; ++q)
            ADV_BLASCHK(ctx, gemm_splitk(blas.h, rocblas_operation_transpose, rocblas_operation_transpose, Ki, Ki, nE,
                                         dV, Ei, 1, dW + size_t(q) * nE * K, Ki, Ki, Cf + size_t(i0 + q) * K * K, Ki,
                                         dsplit));
    }
    if (prec == PSF_C_FP32) {
        k_to_f32<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(Cf, static_cast<float*>(s->dev.C), n);
        ADV_HIPCHK(ctx, hipGetLastError());
    } else if (prec != PSF_C_FP64) {
        ADV_HIPCHK(ctx, launch_pack_c(s, Cf, static_cast<int>(K), st));
    }
    ADV_HIPCHK(ctx, hipStreamSynchronize(st));
    s->has_C = true;
    return PSF_OK;
}

int psf_scene_get_advection(psf_scene* s, double* C)
{
    if (!s) return PSF_ERR_ARG;
    psf_ctx* ctx = s->ctx;
    if (!C) return set_err(ctx, PSF_ERR_ARG, "null argument");
    if (!s->has_C) return set_err(ctx, PSF_ERR_STATE, "advection tensor not uploaded");
    const int prec = s->d.c_precision;
    if (prec != PSF_C_FP64 && prec != PSF_C_FP32)
        return set_err(ctx, PSF_ERR_UNSUPPORTED, "C is stored packed in c_precision %d; only FP64 and FP32 read back", prec);
    const size_t n = size_t(s->d.K) * s->d.K * s->d.K;
    ADV_HIPCHK(ctx, hipSetDevice(ctx->device));
    if (prec == PSF_C_FP64) {
        ADV_HIPCHK(ctx, hipMemcpyAsync(C, s->dev.C, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        ADV_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> t(n);
        ADV_HIPCHK(ctx, hipMemcpyAsync(t.data(), s->dev.C, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        ADV_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t q = 0; q < n; ++q) C[q] = t[q];
    }
    return PSF_OK;
}

}  // extern "C"
