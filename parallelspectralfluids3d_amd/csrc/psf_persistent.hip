// Persistent reduced-step path (MFMA precisions).
#include "psf_persistent.hpp"

namespace psf {

bool persistent_supported(int, int) { return false; }
size_t persistent_c_bytes(int, int) { return 0; }
hipError_t persistent_scene_init(psf_scene*) { return hipSuccess; }
void persistent_scene_free(psf_scene*) {}
int persistent_upload_c(psf_scene*, const double*) { return PSF_ERR_UNSUPPORTED; }
hipError_t launch_persistent_steps(psf_scene*, int, hipStream_t) { return hipErrorNotSupported; }

}  // namespace psf
