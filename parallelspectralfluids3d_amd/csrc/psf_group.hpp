// Basis-shared persistent reduced step (the scene server's batched path):
// one launch runs the reduced steps of several scenes per basis, all of a
// basis's scenes reading one register-resident copy of its C blocks.
#pragma once
#include <cstdint>

#include "psf_common.hpp"

namespace psf {

constexpr int kGroupMaxScenes = 16;   // scenes per instance: two sets of eight B-column pairs
constexpr int kGroupMaxK = 256;       // largest basis K0

// One scene of an instance (device pointers).
struct GroupScene {
    double* omega;        // K: w at the start of the launch; w after its last step on return
    const double* visc;   // K: exp(nu lambda_k dt)
    double* hist;         // w after step j at hist[k * hist_ld + j], j < nsteps (null: not kept)
    int* status;          // set to 1 (non-finite energy) or 2 (exchange timed out)
    double dt;
    int K;                // modes of the scene, K <= K0
    int nsteps;
    int gravity_active;
    int hist_ld;
};

// A basis's C (K0 modes) as BF16 hi+lo blocks in the persistent register
// layout, packed on the device from its FP64 tensor.
struct GroupBasis;
bool group_supported(int K0);
// Cf: K0^3 FP64 on the device, C_i(a, b) at Cf[(i K0 + b) K0 + a]; gravity: K0
// on the device (both stay owned by the caller and must outlive the basis).
// Packs on `stream` and waits for it.
int group_basis_create(const double* Cf, int K0, const double* gravity, hipStream_t stream, GroupBasis** out,
                       std::string* err);
void group_basis_destroy(GroupBasis* b);
int group_basis_workgroups(const GroupBasis* b);   // workgroups of one instance
size_t group_basis_bytes(const GroupBasis* b);     // device bytes of the packed blocks

// Launch tables, pinned staging and exchange records for launches of up to
// max_instances instances.  The staging is rewritten by the next launch:
// the stream must have run the previous launch's copies by then.
struct GroupBatch;
GroupBatch* group_batch_create(int device, int max_instances);
void group_batch_destroy(GroupBatch* b);
int group_batch_capacity(const GroupBatch* b);      // workgroups one launch holds resident
int group_batch_max_instances(const GroupBatch* b);
// The first exchange wait of the batch's launches that gave up (what it
// waited for), cleared on reading; empty when none did.  Synchronous.
std::string group_batch_diag(GroupBatch* b);

struct GroupInstance {
    const GroupBasis* basis;
    int n;                       // scenes, 1..kGroupMaxScenes
    const GroupScene* scenes;    // host array of n
};
// One launch running every instance; instance j uses exchange slot j.
// wait_ticks bounds each exchange wait (s_memrealtime ticks, 0: 2 s).
hipError_t launch_group_steps(GroupBatch* b, const GroupInstance* inst, int n_inst, uint64_t wait_ticks,
                              hipStream_t stream);

}  // namespace psf
