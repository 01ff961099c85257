// Shared host/device definitions of the psf HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "psf/psf.h"

namespace psf {

constexpr int kWave = 64;

// Optional events a launch records itself (hipExtLaunchKernel: the dispatch's
// own start and end, no marker packets between kernels), for timing a launch
// inside a timed region.  Null: an ordinary launch.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};

// Device-side state of one scene.  All arrays live in HBM and stay resident
// between calls; nothing is re-uploaded per frame.
struct DeviceScene {
    // reduced state, FP64
    double* omega = nullptr;     // K
    double* lambda = nullptr;    // K
    double* gravity = nullptr;   // K
    double* visc = nullptr;      // K, exp(nu * lambda_k * dt)
    double* y = nullptr;         // K * plain_splits(K): the plain contraction's partials of y
    int* status = nullptr;       // device status word (1 = non-finite energy)
    // structure tensor
    void* C = nullptr;           // layout and type depend on c_precision
    size_t C_bytes = 0;
    // basis, column-major with leading dimension ldF (zero padded)
    void* Phi = nullptr;
    size_t ldF = 0;
    float* u = nullptr;          // F (velocityField)
    // particles
    float4* parts = nullptr;     // max_particles
    int32_t* cells = nullptr;    // 12 * max_particles (debug output of cell lookups)
    uint32_t* pid = nullptr;     // caller index of each stored particle (storage is in sort_particles' spatial order)
    // frame window (psf_set_frame_window): w after each step of the window,
    // [K][window], u of each of its frames, [window][ldF], and a K-vector for
    // reading one column back
    double* hist = nullptr;
    float* u_hist = nullptr;
    double* omega_view = nullptr;
    // the second buffer pair: the next window is computed into it while the
    // current one is traced
    double* hist_b = nullptr;
    float* u_hist_b = nullptr;
    // particle density (psf_set_density_grid): counters and the half-precision
    // volume, one per cell, padded to a multiple of 2048 cells
    uint32_t* dcounts = nullptr;
    uint16_t* dvol = nullptr;
};

}  // namespace psf

namespace psf {
struct DevicePool;
}

struct psf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::set<psf_scene*> scenes;   // live scenes; destroyed with the context
    std::set<psf_server*> servers; // live servers; destroyed with the context
    psf::DevicePool* pool = nullptr;   // recycled device memory of pooled scenes (servers)
};

struct psf_scene {
    psf_ctx* ctx = nullptr;
    psf_scene_desc d{};
    uint32_t F = 0, V = 0;
    uint32_t nZF = 0, nYF = 0, nXF = 0;
    psf::DeviceScene dev;
    bool has_basis = false, has_phi = false, has_C = false, has_omega = false;
    double dt = 1.0 / 60.0, nu = 0.0;
    int gravity_active = 0;
    float lifetime = 60.0f;
    int path = PSF_PATH_PLAIN;
    int modes_per_wg = 0;        // persistent layout: 0 = one mode per workgroup (up to 256 workgroups)
    uint32_t n_particles = 0;
    uint32_t frame = 0;          // counter of trace passes (respawn draws)
    // frame window: frames computed ahead (filled) and consumed (cursor);
    // while filled > 0, the current w and u are column / row cursor-1 of
    // dev.hist / dev.u_hist and dev.omega holds w at the end of the window
    uint32_t window = 1, win_filled = 0, win_cursor = 0;
    uint32_t win_buf = 0;        // buffer of the frames being consumed (0: hist/u_hist, 1: hist_b/u_hist_b)
    bool win_pipeline = true;    // compute the next window on the side stream during this one
    bool win_ahead = false;      // the next window is queued on the side stream into the other buffer
    // particle passes of frames already advanced but not yet launched: rows
    // trace_c0 .. trace_c0 + trace_pending - 1 of the current window buffer
    // (they run as one launch; every read of the particles launches them)
    uint32_t trace_pending = 0, trace_c0 = 0;
    uint32_t win_ahead_n = 0;    // frames in the window queued on the side stream
    // a bounded run (psf_time_phase's frame phases): window fills compute at
    // most fill_budget more frames, so nothing past the run is computed ahead
    bool bounded = false;
    uint32_t fill_budget = 0;
    psf_scene_counters cnt{};    // psf_scene_get_counters
    // kernels of a timed region (psf_time_phase): an event pair around each
    // step, velocity-field and particle launch, summed by kind at the end
    struct TimedLaunch {
        int kind;                // 0 steps, 1 velocity fields, 2 particle frames
        uint32_t units;
        hipEvent_t a, b;
    };
    bool time_kernels = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> ev_pool;   // events of timed launches, reused across regions
    size_t ev_used = 0;
    psf_region_kernels region{};       // psf_scene_region_kernels
    hipStream_t side = nullptr;
    hipEvent_t ev_ready[2] = {nullptr, nullptr};   // side stream: buffer filled
    hipEvent_t ev_fork = nullptr;                  // context stream position the side stream starts from
    uint32_t seed = 0;
    void* host = nullptr;        // host-side copies (psf_scene_host)
    void* persistent = nullptr;  // persistent-path state (PersistentState)
    bool pooled = false;         // device memory comes from ctx->pool and goes back there
    bool phi_borrowed = false;   // dev.Phi belongs to someone else (a server's shared basis)
    // uploads return without waiting for the stream (a server's scenes: their
    // host sources outlive the copies, and the first launch waits on the
    // upload stream through an event)
    bool async_uploads = false;
    // particle density grid: cells per axis (0: none), cell count rounded up
    // to a multiple of 2048, whether particle passes count their new positions,
    // and what the counters hold: 0 zeros, 1 the counts of the stored
    // particles, 2 anything else
    uint32_t dens_n[3] = {0, 0, 0};
    size_t dens_cells = 0;
    bool dens_in_trace = false;
    int dens_state = 0;
};

namespace psf {
// device memory of one scene: hipMalloc / hipFree, or the context's pool for
// pooled scenes (size classes recycled without hipFree, which synchronises)
hipError_t scene_alloc(psf_scene* s, void** p, size_t bytes);
void scene_free(psf_scene* s, void* p);
}  // namespace psf
