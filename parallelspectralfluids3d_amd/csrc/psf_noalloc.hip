// psf_step_device: plain reduced steps on caller-owned device buffers,
// allocating nothing (include/psf/psf.h).  The kernels are the plain path's
// (k_contract_plain, k_update); the partial sums of y that a scene keeps in
// its own buffer live here in one file-scope device array per GPU.
#include <hip/hip_runtime.h>

#include <mutex>

#include "psf_kernels.hpp"

namespace psf {
namespace {

// largest K, and so the partials of K * plain_splits(K) doubles: 256 KiB
constexpr int kNoAllocMaxK = 1024;
constexpr int kNoAllocParts = kNoAllocMaxK * ((kNoAllocMaxK + 31) / 32);
constexpr int kMaxDevices = 64;

// Host code never uses this name as an address: each GPU holds its own copy,
// whose address comes from hipGetSymbolAddress with that GPU current.
__device__ __align__(256) double scratch[kNoAllocParts];

std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};

// this GPU's copy of scratch, looked up once per device with it current
hipError_t scratch_on(int device, double** out)
{
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace
}  // namespace psf

using namespace psf;

extern "C" int psf_step_device(int c_precision, const void* C, double* omega, const double* visc,
                               const double* gravity, uint32_t K, double dt, int gravity_active, int nsteps,
                               int* status, void* stream)
{
    if (c_precision != PSF_C_FP64 && c_precision != PSF_C_FP32) return PSF_ERR_UNSUPPORTED;
    if (K < 1 || K > static_cast<uint32_t>(kNoAllocMaxK) || nsteps < 0) return PSF_ERR_ARG;
    if (!C || !omega || !visc || !status || (gravity_active && !gravity)) return PSF_ERR_ARG;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return PSF_ERR_HIP;
    double* part = nullptr;
    if (scratch_on(device, &part) != hipSuccess) return PSF_ERR_HIP;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int k = static_cast<int>(K);
    for (int it = 0; it < nsteps; ++it) {
        if (launch_contract_plain(c_precision, C, omega, part, k, st) != hipSuccess) return PSF_ERR_HIP;
        if (launch_update(omega, part, visc, gravity, k, dt, gravity_active ? 1 : 0, status, nullptr, 0, 0, st) !=
            hipSuccess)
            return PSF_ERR_HIP;
    }
    return PSF_OK;
}
