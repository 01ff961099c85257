// Persistent reduced-step path (MFMA precisions): one launch runs n steps.
#pragma once
#include <vector>

#include "psf_common.hpp"

namespace psf {

// ipw_req: modes per workgroup (0 = one per workgroup, up to 256 workgroups)
bool persistent_supported(int c_precision, int K, int ipw_req = 0);
size_t persistent_c_bytes(int c_precision, int K, int ipw_req = 0);
int persistent_workgroups(int c_precision, int K, int ipw_req = 0);
// largest modes-per-workgroup (<= 8) whose blocks all stay in registers
int persistent_auto_ipw(int c_precision, int K);
hipError_t persistent_scene_init(psf_scene* s);
void persistent_scene_free(psf_scene* s);
int persistent_upload_c(psf_scene* s, const double* C);
// bound of each wait of the persistent step's exchange, in s_memrealtime
// ticks (100 MHz); 0 restores the default of 2 s
int persistent_set_wait_ticks(psf_scene* s, uint64_t ticks);
// hist != nullptr also stores w after step j at hist[k * hist_ld + j]
hipError_t launch_persistent_steps(psf_scene* s, int nsteps, hipStream_t st, double* hist = nullptr,
                                   int hist_ld = 0, const LaunchEvents* ev = nullptr);

// Several scenes (all BF16 / BF16 hi+lo, or all FP16) in one launch; scene j
// runs nsteps[j] steps.  The tables go through pinned staging buffers that
// the next call overwrites: synchronise the stream before reusing the batch.
struct PersistentBatch;
PersistentBatch* persistent_batch_create(int device, int max_scenes, int max_wgs);
void persistent_batch_destroy(PersistentBatch* b);
int persistent_batch_wg_capacity(const PersistentBatch* b, int c_precision);
hipError_t launch_persistent_batch(PersistentBatch* b, psf_scene* const* sc, const int* nsteps, double* const* hist,
                                   const int* hist_ld, int n, hipStream_t st);

// Packs C of the scene's K modes (BF16 or BF16 hi+lo) on the device from an
// FP64 tensor of K0 >= K modes, C_i(a, b) at Cf[(i * K0 + b) * K0 + a].
hipError_t launch_pack_c(psf_scene* s, const double* Cf, int K0, hipStream_t st);

// The block layout of K modes at ipw_req modes per workgroup (0: one per
// workgroup): G workgroups of bpw blocks, ipw modes each, and the metadata
// word of each block (t << 24 | row tile << 16 | q1 << 8 | q2).
void persistent_layout_meta(int c_precision, int K, int ipw_req, int& G, int& ipw, int& bpw,
                            std::vector<uint32_t>& meta);
// Packs the first K modes of the FP64 tensor Cf (K0 modes) into that layout
// (hilo: BF16 hi+lo blocks, else BF16); inv_scale receives G * ipw ones.
hipError_t launch_pack_c_layout(const double* Cf, int K0, int K, int G, int ipw, int bpw, int hilo,
                                const uint32_t* meta, uint4* blocks, double* inv_scale, hipStream_t st);

}  // namespace psf
