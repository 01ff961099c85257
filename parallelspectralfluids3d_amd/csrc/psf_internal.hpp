// Library-internal entry points shared between the C ABI (psf_api.hip) and
// the scene server (psf_server.hip).
#pragma once
#include <string>
#include <vector>

#include "psf_common.hpp"
#include "psf_kernels.hpp"

namespace psf {

struct SceneOptions {
    int modes_per_wg = 0;          // persistent layout (0 = one mode per workgroup)
    void* borrowed_phi = nullptr;  // use this Phi (F x K' column-major, K' >= K, leading dim ldF) instead of a copy
    bool cells = true;             // allocate the cell-lookup debug buffer
    bool pooled = false;           // take device memory from the context's pool
    bool frame_pipeline = true;    // frame windows keep a second buffer pair and a side stream
    bool async_uploads = false;    // set-up calls do not wait for their copies (host sources must outlive them)
    bool no_c = false;             // no structure tensor of its own: its steps run from a shared basis's C
    bool no_init = false;          // no fill kernels: the caller writes omega, lambda, gravity, visc and status
};

int scene_create_ex(psf_ctx* ctx, const psf_scene_desc* desc, const SceneOptions& opt, psf_scene** out);
// the spatial storage order of n particles (psf_set_particles): sorted[j] = xyzw[pid[j]]
void sort_particles(const psf_scene_desc& d, const float* xyzw, uint32_t n, float* sorted, uint32_t* pid);
// Host-side state of a scene whose device arrays the caller fills with its
// own copies: particles in storage order with caller indices pid, lambda,
// the step parameters, omega present.  Nothing is uploaded.
// caller index of each stored particle of a scene
const std::vector<uint32_t>& scene_pid(const psf_scene* s);
void scene_adopt_state(psf_scene* s, std::vector<uint32_t>&& pid, const double* lambda, double dt, double nu,
                       int gravity_active, float lifetime);
// destroy a scene whose device work is complete (no wait on the context stream)
void scene_release(psf_scene* s);
// omega and particles (caller's order) by synchronous null-stream copies
int scene_read_results(psf_scene* s, double* omega, float* xyzw);
TraceParams scene_trace_params(const psf_scene* s);
// launches the particle passes of deferred frames
int scene_flush_traces(psf_scene* s);
// ends a frame window early (the state-changing calls' first step)
int scene_window_flush(psf_scene* s);
int scene_check_status(psf_scene* s);
int ctx_set_error(psf_ctx* ctx, int code, const std::string& msg);
size_t pool_class_bytes(size_t bytes);   // bytes a pooled allocation of this size holds
// the density grid a particle pass counts into (counts null: none); clears
// the counters first when they hold anything
int density_for_trace(psf_scene* s, DensityGrid& D);
// the stored particles changed without being counted
void density_particles_changed(psf_scene* s);
void density_release(psf_scene* s);
int density_enqueue_count(psf_scene* s);    // count pass of the stored particles
int density_enqueue_volume(psf_scene* s);   // volume pass (clears the counters)
// psf_density on the context stream; counts_host (optional) receives the
// counts by an asynchronous copy
int density_enqueue(psf_scene* s, uint32_t* counts_host);

}  // namespace psf
