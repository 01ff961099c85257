// Launchers of the psf HIP kernels (implemented in psf_kernels.hip).
#pragma once
#include "psf_common.hpp"

namespace psf {

struct TraceParams {
    uint32_t nx, ny, nz;
    uint32_t nZF, nYF;          // face-block offsets (Z faces first, then Y, then X)
    float aabb_min[3], aabb_max[3];
    float ext[3];               // aabb_max - aabb_min (AABB::getExtent, float)
    double res;                 // AbstractSolver::resolution
    double inv_res;             // RN(1 / res)
    float hf;                   // (float)resolution
    float inv_h;                // float(1.0 / resolution)
    float dtf;                  // float(timeStep)
    float lifetime;
    float center_z;             // AABB::getCenter().z (float)
    uint32_t seed, frame;
};

// one scene's share of a batched window reconstruction (PHI_FP32)
struct WindowScene {
    const void* Phi;
    size_t ldF;
    uint32_t F;
    int K;
    const double* hist;         // [K][window]
    float* uh;                  // [window][ldF]
    uint32_t block0;            // first block of this scene in the launch
};

// one scene's share of a batched multi-frame particle pass
struct TraceScene {
    TraceParams P;              // P.frame = frame of the first of the nframes
    const float* u_hist;        // u of frame t at u_hist + t * ldU
    size_t ldU;
    float4* parts;
    const uint32_t* pid;
    uint32_t n, nframes;
    uint32_t block0;
};

// reduced step, plain path: y_i = w^T C_i w (one block per mode i), then the
// FP64 update of w (one block)
hipError_t launch_contract_plain(int c_precision, const void* C, const double* omega, double* y, int K,
                                 hipStream_t s);
// hist != nullptr also stores the new w at hist[k * hist_ld + hist_t]
hipError_t launch_update(double* omega, const double* y, const double* visc, const double* gravity, int K,
                         double dt, int gravity_active, int* status, double* hist, int hist_ld, int hist_t,
                         hipStream_t s);
// u = float(Phi w)
hipError_t launch_reconstruct(int phi_precision, const void* Phi, size_t ldF, uint32_t F, int K,
                              const double* omega, float* u, hipStream_t s);
// u_t = float(Phi w_t) for t < window (2, 4, 8, 16 or 32), w_t = hist[k * window + t],
// u_t at uh + t * ldF; bitwise equal to launch_reconstruct for each w_t
hipError_t launch_reconstruct_window(int phi_precision, int window, const void* Phi, size_t ldF, uint32_t F, int K,
                                     const double* hist, float* uh, hipStream_t s);
// all scenes of d_tab (device memory, ascending block0) in one launch of
// `blocks` blocks; window_blocks() gives one scene's block count
hipError_t launch_reconstruct_window_multi(int phi_precision, int window, const WindowScene* d_tab, int n,
                                           uint32_t blocks, hipStream_t s);
uint32_t window_blocks(int window, size_t ldF);
// nframes particle passes of every scene of d_tab in one launch (256 particles per block)
hipError_t launch_trace_window_multi(const TraceScene* d_tab, int n, uint32_t blocks, hipStream_t s);
// out[k] = hist[k * ld + t]
hipError_t launch_gather_column(const double* hist, int ld, int t, int K, double* out, hipStream_t s);
// one particle pass
hipError_t launch_trace(const TraceParams& p, const float* u, float4* parts, const uint32_t* pid, uint32_t n,
                        int32_t* cells, hipStream_t s);
hipError_t launch_zero_int(int* p, int n, hipStream_t s);

}  // namespace psf
