// Persistent reduced-step path (MFMA precisions): one launch runs n steps.
#pragma once
#include "psf_common.hpp"

namespace psf {

bool persistent_supported(int c_precision, int K);
size_t persistent_c_bytes(int c_precision, int K);
hipError_t persistent_scene_init(psf_scene* s);
void persistent_scene_free(psf_scene* s);
int persistent_upload_c(psf_scene* s, const double* C);
// hist != nullptr also stores w after step j at hist[k * hist_ld + j]
hipError_t launch_persistent_steps(psf_scene* s, int nsteps, hipStream_t st, double* hist = nullptr,
                                   int hist_ld = 0);

}  // namespace psf
