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
    uint32_t pow2;              // resolution is 2^-k (so inv_h = 2^k exactly)
};

// Particle density grid: counts of particles per cell of an n[0] x n[1] x n[2]
// grid spanning the AABB, x fastest, then y, then z (a 3-D image's texel
// order).  Counting is off when counts is null.
struct DensityGrid {
    float amin[3], ext[3];      // AABB min and AABB::getExtent (float)
    float r[3];                 // cells per axis as float (exact: n < 2^24)
    uint32_t n[3];              // cells per axis
    uint32_t* counts;           // n[0]*n[1]*n[2] counters
};

// Cell of a coordinate on one axis: trunc((p - amin) / ext * r) in float,
// clamped to [0, n-1]; NaN lands in cell 0.  Division and product are
// IEEE-rounded (no contraction, correctly rounded division).
__device__ __forceinline__ uint32_t density_axis(float p, float amin, float ext, float r, uint32_t n)
{
#pragma clang fp contract(off)
    const float t = __fdiv_rn(p - amin, ext) * r;
    if (t >= r) return n - 1u;
    if (t > 0.0f) return static_cast<uint32_t>(t);
    return 0u;
}

__device__ __forceinline__ uint32_t density_cell(const DensityGrid& D, float x, float y, float z)
{
    const uint32_t cx = density_axis(x, D.amin[0], D.ext[0], D.r[0], D.n[0]);
    const uint32_t cy = density_axis(y, D.amin[1], D.ext[1], D.r[1], D.n[1]);
    const uint32_t cz = density_axis(z, D.amin[2], D.ext[2], D.r[2], D.n[2]);
    return (cz * D.n[1] + cy) * D.n[0] + cx;
}

// one scene's share of a batched window reconstruction (PHI_FP32)
struct WindowScene {
    const void* Phi;
    size_t ldF;
    uint32_t F;
    int K;
    const double* hist;         // [K][window]
    float* uh;                  // [window][ldF]
    uint32_t block0;            // first block of this entry in the launch
    int K2;                     // a second scene on the same Phi with K2 <= K modes (0: none; FP32 sums only)
    const double* hist2;
    float* uh2;
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

// reduced step, plain path: partials of y_i = w^T C_i w over column chunks
// (plain_splits(K) per mode, K * plain_splits(K) doubles in part), then the
// FP64 update of w (one block), which sums them
int plain_splits(int K);
hipError_t launch_contract_plain(int c_precision, const void* C, const double* omega, double* part, int K,
                                 hipStream_t s);
// hist != nullptr also stores the new w at hist[k * hist_ld + hist_t]
hipError_t launch_update(double* omega, const double* part, const double* visc, const double* gravity, int K,
                         double dt, int gravity_active, int* status, double* hist, int hist_ld, int hist_t,
                         hipStream_t s);
// u = float(Phi w)
hipError_t launch_reconstruct(int phi_precision, const void* Phi, size_t ldF, uint32_t F, int K,
                              const double* omega, float* u, hipStream_t s, const LaunchEvents* ev = nullptr);
// u_t = float(Phi w_t) for t < ncols (1..32), w_t = hist[k * hist_ld + t],
// u_t at uh + t * ldF; bitwise equal to launch_reconstruct for each w_t
hipError_t launch_reconstruct_window(int phi_precision, int hist_ld, int ncols, const void* Phi, size_t ldF, uint32_t F,
                                     int K, const double* hist, float* uh, hipStream_t s, const LaunchEvents* ev = nullptr);
// all scenes of d_tab (device memory, ascending block0) in one launch of
// `blocks` blocks; window_blocks() gives one scene's block count
hipError_t launch_reconstruct_window_multi(int phi_precision, int window, const WindowScene* d_tab, int n,
                                           uint32_t blocks, hipStream_t s);
uint32_t window_blocks(int window, size_t ldF);
// nframes particle passes of every scene of d_tab in one launch (256 particles per block)
hipError_t launch_trace_window_multi(const TraceScene* d_tab, int n, uint32_t blocks, hipStream_t s);
// out[k] = hist[k * ld + t]
hipError_t launch_gather_column(const double* hist, int ld, int t, int K, double* out, hipStream_t s);
// one particle pass; with D.counts set, each particle's new position is also
// counted into D's grid
hipError_t launch_trace(const TraceParams& p, const float* u, float4* parts, const uint32_t* pid, uint32_t n,
                        int32_t* cells, const DensityGrid& D, hipStream_t s, const LaunchEvents* ev = nullptr);
// nframes particle passes in one launch, frame t with u0 + t * ldU and respawn
// frame p.frame + t; D counts the positions after the last
hipError_t launch_trace_frames(const TraceParams& p, const float* u0, size_t ldU, uint32_t nframes, float4* parts,
                               const uint32_t* pid, uint32_t n, const DensityGrid& D, hipStream_t s, const LaunchEvents* ev = nullptr);
// counts of the n stored particles into D's grid (added to what it holds)
hipError_t launch_density_count(const DensityGrid& D, const float4* parts, uint32_t n, hipStream_t s);
// volume[v] = half(float(counts[v]) / 1e6) and counts[v] = 0, for v < n2048 * 2048
hipError_t launch_density_volume(uint32_t* counts, uint16_t* volume, size_t n2048, hipStream_t s);
hipError_t launch_zero_int(int* p, int n, hipStream_t s);
// k_timed_region_mark: an empty launch (which = 0 start, 1 end of a timed region)
hipError_t launch_timed_region_mark(int which, hipStream_t s);
hipError_t launch_fill_f64(double* p, int n, double v, hipStream_t s);

}  // namespace psf
