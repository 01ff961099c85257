/*
 * psf.h -- C ABI of the MI355X reduced-space fluid step (libpsf_hip.so).
 *
 * The library replaces the per-frame hot path of the reference solver
 * (Solver/SolverImpl/psfsolver.cpp:21-234 on the CPU and
 *  Solver/SolverImpl/psfsolvergpu.cpp:216-414 + Resources/Simulation/simulation.cl
 *  on OpenCL):
 *   1. the reduced advection step  y_k = w^T C_k w, w += dt*y, energy
 *      renormalisation, viscous decay, forcing       (psfsolver.cpp:34-59)
 *   2. the face-flux reconstruction u = float(Phi w)  (psfsolver.cpp:60)
 *   3. particle tracing by trilinear face-flux interpolation
 *                                                     (psfsolver.cpp:70-232)
 * Preprocessing (mesh, eigenbasis, structure tensor C) stays with the caller,
 * as in the reference, where AbstractSolver's protected members are the data
 * contract between buildLaplace/buildAdvection and integrate()
 * (Solver/abstractsolver.h:94-126).
 *
 * Conventions
 *  - every function returns 0 (PSF_OK) or a negative psf_status; the message
 *    of the last failure on a context is psf_last_error(ctx).  Nothing exits.
 *  - host arrays passed in are copied; device memory is owned by the library.
 *  - work is queued on the context's stream; functions that return data to
 *    the host synchronise that stream.  A context is used by one thread at a
 *    time.
 *  - array layouts are the reference's Eigen layouts (column-major).
 */
#ifndef PSF_PSF_H
#define PSF_PSF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSF_ABI_VERSION 4

#if defined(__GNUC__)
#define PSF_API __attribute__((visibility("default")))
#else
#define PSF_API
#endif

typedef enum psf_status {
    PSF_OK = 0,
    PSF_ERR_ARG = -1,          /* invalid argument or shape */
    PSF_ERR_HIP = -2,          /* HIP runtime failure */
    PSF_ERR_NOMEM = -3,        /* device or host allocation failed */
    PSF_ERR_STATE = -4,        /* call out of order (e.g. step before upload) */
    PSF_ERR_UNSUPPORTED = -5,  /* combination not built (e.g. path/precision) */
    PSF_ERR_NONFINITE = -6     /* the FP64 energy check saw 0, NaN or Inf */
} psf_status;

/* storage/arithmetic of the structure tensor C in the reduced step */
typedef enum psf_c_precision {
    PSF_C_FP64 = 0,    /* FP64 C, FP64 arithmetic (parity path) */
    PSF_C_FP32 = 1,    /* FP32 C, FP32 products, FP64 reduction */
    PSF_C_BF16 = 2,    /* BF16 C on MFMA, FP32 accumulation */
    PSF_C_FP16 = 3,    /* FP16 C (scaled per mode) on MFMA, FP32 accumulation */
    PSF_C_BF16X2 = 4   /* C = hi + lo, both BF16, on MFMA, FP32 accumulation */
} psf_c_precision;

/* storage of the basis Phi and arithmetic of u_f = sum_k Phi[f][k] w_k */
typedef enum psf_phi_precision {
    PSF_PHI_FP64 = 0,       /* FP64 Phi, FP64 sums (parity path) */
    PSF_PHI_FP32 = 1,       /* FP32 Phi, FP64 sums */
    PSF_PHI_FP32_ACC32 = 2  /* FP32 Phi and w, FP32 sums on the FP32 matrix cores */
} psf_phi_precision;

/* how the reduced step is launched */
typedef enum psf_step_path {
    PSF_PATH_PLAIN = 0,       /* two launches per step: contraction, FP64 update */
    PSF_PATH_PERSISTENT = 1   /* one launch for n steps, C kept on chip, one
                                 device-wide exchange of y per step (MFMA modes) */
} psf_step_path;

typedef struct psf_ctx psf_ctx;
typedef struct psf_scene psf_scene;
typedef struct psf_server psf_server;

/* Grid and basis description.  Counts follow DECMesh3D (DEC/decmesh3d.cpp:64-70):
   F = (nz+1)nx ny + (ny+1)nx nz + (nx+1)ny nz faces, V = nx ny nz voxels.
   Limits: 1 <= nx, ny, nz <= 2048 and F < 2^30. */
typedef struct psf_scene_desc {
    uint32_t nx, ny, nz;       /* voxel counts (DECMesh3D::getDimensions) */
    double resolution;         /* AbstractSolver::resolution, the voxel size h */
    float aabb_min[3];         /* Model::getAABB().min; the particle cell lookup
                                  (psfsolver.cpp:80-89) assumes max == -min */
    float aabb_max[3];
    uint32_t K;                /* AbstractSolver::nEigenFunctions */
    int32_t c_precision;       /* psf_c_precision */
    int32_t phi_precision;     /* psf_phi_precision */
    uint32_t max_particles;    /* AbstractSolver::maxParticles */
} psf_scene_desc;

/* ---- context --------------------------------------------------------------- */
PSF_API int psf_ctx_create(int device, psf_ctx** out);
/* also destroys every scene still alive on the context; their handles become invalid */
PSF_API int psf_ctx_destroy(psf_ctx* ctx);
PSF_API const char* psf_last_error(const psf_ctx* ctx);
PSF_API int psf_ctx_sync(psf_ctx* ctx);
/* the context's hipStream_t, for callers that order their own work after ours */
PSF_API void* psf_ctx_stream(psf_ctx* ctx);
PSF_API int psf_abi_version(void);

/* ---- scene lifetime and uploads -------------------------------------------
 * Replaces PSFSolverGPU::buildAdvection's upload tail
 * (Solver/SolverImpl/psfsolvergpu.cpp:718-874). */
PSF_API int psf_scene_create(psf_ctx* ctx, const psf_scene_desc* desc, psf_scene** out);
PSF_API int psf_scene_destroy(psf_scene* scene);
/* bytes of device memory the scene holds (admission by footprint) */
PSF_API int psf_scene_footprint(const psf_scene_desc* desc, uint64_t* bytes);

/* Phi: F x K column-major (AbstractSolver::velBasisField), or NULL for a
   scene that only runs reduced steps; lambda: K (eigenValues); gravity: K
   (gravity = Phi^T f, psfsolver.cpp:344-358), or NULL for zero. */
PSF_API int psf_scene_upload_basis(psf_scene* scene, const double* Phi, const double* lambda,
                           const double* gravity);
/* C: K matrices K x K, each column-major, matrix i at C + i*K*K
   (std::vector<Eigen::MatrixXd> advection, abstractsolver.h:118). */
PSF_API int psf_scene_upload_advection(psf_scene* scene, const double* C);
/* The same tensor as K separate matrices: mats[i] points at the K x K
   column-major C_i, i.e. advection[i].data() of the reference's
   std::vector<Eigen::MatrixXd> (abstractsolver.h:118). */
PSF_API int psf_scene_upload_advection_matrices(psf_scene* scene, const double* const* mats);
/* Optional: the caller's DECMesh3D::getSignBitString() (decmesh3d.cpp:82-106,
   1390-1399).  The tracer evaluates the sign bits in closed form; this call
   checks the caller's table against that closed form and fails on mismatch. */
PSF_API int psf_scene_check_signs(psf_scene* scene, const uint32_t* words, size_t nwords);

/* ---- closed boxes: basis and C built on the device ---------------------------
 * No counterpart in the reference, which eigensolves M (psfsolver.cpp:236-342)
 * and sums C over voxels on the CPU (psfsolver.cpp:431-551).  For a closed box
 * (every voxel inside) the discrete sin/cos modes form a divergence-free basis
 * with closed forms, and C factorises into 1-D sums along x, y and z, so both
 * are built in place on the device from the mode table: nothing of size F or E
 * crosses PCIe and C takes O(K^3) work.
 * freqs: K rows (a, b, c) of frequencies along x, y, z (0 <= a < nx, ...);
 * amps: K rows of the x-, y- and z-face amplitudes of the unit-norm mode;
 * lambda: K eigenvalues; gravity: K or NULL (as for psf_scene_upload_basis).
 * Mode k's flux on an X face (x node, y cell, z cell) is
 *   amps[3k] * sin(a pi x / nx) * cos(b pi (y+1/2) / ny) * cos(c pi (z+1/2) / nz),
 * and likewise on Y and Z faces with the sine along the face's own axis. */
PSF_API int psf_scene_generate_box_basis(psf_scene* scene, const int32_t* freqs, const double* amps,
                                         const double* lambda, const double* gravity);
/* C of the same modes in the scene's c_precision (not FP16, whose per-mode
   scales need the host upload). */
PSF_API int psf_scene_build_box_advection(psf_scene* scene, const int32_t* freqs, const double* amps,
                                          const double* lambda);

/* setTimestep / setViscosity / setGravityActive / setLifeTime
   (Solver/abstractsolver.cpp:226-250).  lifetime is the value stored in
   AbstractSolver::lifeTime (frames). */
PSF_API int psf_set_params(psf_scene* scene, double dt, double viscosity, int gravity_active, float lifetime);
PSF_API int psf_set_step_path(psf_scene* scene, int path);
/* Watchdog of the persistent path (no counterpart in the reference): every
   wait of a workgroup for the others' y in the device-wide exchange gives up
   after `seconds` (default 2 s; 0 restores it), every workgroup then leaves
   the launch, and the next call that reads results returns PSF_ERR_HIP
   ("exchange timed out") instead of hanging.  omega keeps the last step all
   workgroups completed.  PSF_ERR_UNSUPPORTED for plain-path scenes. */
PSF_API int psf_set_wait_limit(psf_scene* scene, double seconds);

/* basisCoeff (AbstractSolver::basisCoeff) */
PSF_API int psf_set_omega(psf_scene* scene, const double* omega);
PSF_API int psf_get_omega(psf_scene* scene, double* omega);

/* particles: n records (x, y, z, w) as glm::vec4 (Graphics/Particle/Particle.h:14).
   The library stores them sorted by containing voxel; get/trace return the
   caller's order, and respawn draws use the caller's particle index.
   psf_get_particles must ask for exactly the count that was set. */
PSF_API int psf_set_particles(psf_scene* scene, const float* xyzw, uint32_t n);
PSF_API int psf_get_particles(psf_scene* scene, float* xyzw, uint32_t n);

/* ---- the hot path ----------------------------------------------------------- */
/* nsteps reduced steps (psfsolver.cpp:34-59) on the scene's path */
PSF_API int psf_step(psf_scene* scene, int nsteps);
/* u = float(Phi w) on the device (psfsolver.cpp:60) */
PSF_API int psf_reconstruct(psf_scene* scene);
/* velocityField (AbstractSolver::velocityField), F floats */
PSF_API int psf_get_velocity(psf_scene* scene, float* u);
/* one pass of the particle loop with the current u (psfsolver.cpp:70-232);
   cells (optional, host, 12 ints per particle) receives the clamped
   xOfs,yOfs,zOfs,xOfs-,xOfs+,yOfs-,yOfs+,zOfs-,zOfs+ of advected particles
   (-1 for respawned ones) */
PSF_API int psf_trace(psf_scene* scene, int32_t* cells);
/* nframes x AbstractSolver::integrate(): step, reconstruct, trace.  With a
   frame window the particle passes of consecutive frames are deferred and
   run up to 8 frames per launch (the same bits as one pass per frame); every
   call that reads or replaces the particles, psf_density and psf_ctx_sync
   launch the deferred passes first. */
PSF_API int psf_integrate(psf_scene* scene, int nframes);
/* Frame window (no counterpart in the reference; a scheduling choice inside
   integrate()).  With frames = W > 1, when integrate() finds no frame
   computed ahead it runs the next W reduced steps in one launch and the W
   reconstructions u_t = float(Phi w_t) in one pass over Phi, then each
   integrate() traces with its own u_t.  The reduced dynamics do not depend
   on the particles, so omega, u and the particles are bitwise those of W
   single frames.  Any call that changes the scene (params, omega, uploads,
   path, psf_step, psf_reconstruct) first drops the frames computed ahead.
   W is 1 (default: one frame at a time), 2, 4, 8, 16 or 32; the window
   holds W * (4 * F' + 8 * K) + 8 * K more bytes, F' = F rounded up to 256. */
PSF_API int psf_set_frame_window(psf_scene* scene, uint32_t frames);

/* What a scene's calls have queued since it was created (no counterpart in
   the reference; for tests and measurement).  A window fill is one launch of
   reduced steps plus one reconstruction of their velocity fields. */
typedef struct psf_scene_counters {
    uint64_t frames;             /* integrate() frames queued */
    uint64_t window_fills;       /* frame windows computed */
    uint64_t window_frames;      /* frames those fills computed */
    uint64_t windows_dropped;    /* windows ended early by a state change */
    uint64_t frames_dropped;     /* frames computed ahead and discarded by those drops */
    uint64_t step_launches;      /* reduced-step launches (a persistent launch, or one plain step) */
    uint64_t recon_launches;     /* velocity-field launches (one frame or one window) */
    uint64_t particle_launches;  /* particle-pass launches (one or several frames each) */
} psf_scene_counters;
PSF_API int psf_scene_get_counters(psf_scene* scene, psf_scene_counters* out);

/* Kernel time inside the last timed frame region of psf_time_phase (phase
   11; no counterpart in the reference; for measurement): every persistent
   reduced-step launch, velocity-field launch and particle-pass launch of the
   region records its own start and end (the dispatch's timestamps, no extra
   packets between the kernels), and the durations are summed by kind.  The
   plain path's reduced steps are not timed.  units: reduced steps, velocity
   fields, particle frames. */
typedef struct psf_region_kernels {
    double step_ms, recon_ms, particle_ms;
    uint32_t step_launches, recon_launches, particle_launches;
    uint32_t steps, recon_frames, particle_frames;
} psf_region_kernels;
PSF_API int psf_scene_region_kernels(psf_scene* scene, psf_region_kernels* out);

/* Device addresses of a scene's resident buffers, for a caller's own GPU
   code (a renderer drawing the particles or u in place, as the reference's
   OpenCL solver shares its particle buffer with OpenGL,
   psfsolvergpu.cpp:358-414).  Valid until the scene changes shape or is
   destroyed; u and the particle records are those of the current frame once
   psf_ctx_sync has returned (deferred particle passes run first).  Layouts
   as for the uploads (C only for FP64/FP32 scenes, else NULL); particle
   records are stored in a spatial order, parts[j] holding caller particle pid[j]. */
typedef struct psf_scene_view {
    double* omega;               /* K */
    const double* lambda;        /* K */
    const double* gravity;       /* K */
    const double* visc;          /* K: exp(viscosity * lambda_k * dt) */
    const void* C;               /* K matrices K x K, column-major, FP64 or FP32; NULL for packed precisions */
    const void* Phi;             /* column-major, leading dimension ld_phi, FP64 or FP32 */
    uint64_t ld_phi;
    const float* u;              /* F */
    float* particles;            /* n_particles x (x, y, z, w) */
    const uint32_t* pid;         /* n_particles */
    uint32_t n_particles;
    int* status;                 /* 1 after a zero or non-finite energy */
} psf_scene_view;
PSF_API int psf_scene_device_view(psf_scene* scene, psf_scene_view* out);

/* Reduced steps on caller-owned device buffers, allocating nothing.  No
   counterpart in the reference; it is for callers that cannot allocate
   device memory (code under HIP graph capture, serving loops with
   preallocated buffers).  nsteps steps of the plain path of psf_step
   (psfsolver.cpp:34-59) on device arrays: C (K matrices K x K, column-major,
   PSF_C_FP64 or PSF_C_FP32), omega (K, updated in place), visc (K:
   exp(viscosity * lambda_k * dt)), gravity (K; may be NULL when
   gravity_active is 0) and status (one int, set to 1 on a zero or
   non-finite energy, never cleared here).  Runs on the current device, queued
   on `stream` (a hipStream_t of that device; NULL: the null stream), and
   returns without waiting.  The partial sums of y live in one library-owned
   device array per GPU (256 KiB, K <= 1024), held until the process ends:
   calls that overlap on one GPU race, so order them on one stream.
   PSF_ERR_ARG for K outside [1, 1024], nsteps < 0 or a null pointer;
   PSF_ERR_UNSUPPORTED for other precisions.  psf_scene_device_view gives a
   scene's buffers in these layouts. */
PSF_API int psf_step_device(int c_precision, const void* C, double* omega, const double* visc,
                            const double* gravity, uint32_t K, double dt, int gravity_active, int nsteps,
                            int* status, void* stream);

/* ---- scene cache -----------------------------------------------------------------
 * A versioned file of a scene's precomputed inputs: Phi (optional, FP64, F x K
 * column-major), lambda, g (optional) and C (FP64, K matrices K x K), i.e. what
 * PSFSolver::buildLaplace / buildEigenFunctions / buildAdvection produce
 * (Solver/SolverImpl/psfsolver.cpp:236-551), with the grid, resolution, AABB and
 * K.  No counterpart in the reference, which rebuilds them on every start.
 * The header carries a format version and an FNV-1a checksum of the payload;
 * readers refuse other versions (PSF_ERR_UNSUPPORTED), truncated or corrupt
 * files (PSF_ERR_ARG).  Writes go to path.tmp, then rename.
 * psf_cache_write / psf_cache_read_header / psf_cache_read touch no GPU.
 * psf_scene_load_cache uploads a cache into a scene created with the same
 * grid, resolution and K, through psf_scene_upload_basis and
 * psf_scene_upload_advection, in the scene's precisions. */
PSF_API int psf_cache_write(const char* path, const psf_scene_desc* desc, const double* Phi, const double* lambda,
                            const double* gravity, const double* C);
PSF_API int psf_cache_read_header(const char* path, psf_scene_desc* desc, uint32_t* has_phi, uint32_t* has_gravity);
PSF_API int psf_cache_read(const char* path, double* Phi, double* lambda, double* gravity, double* C);
PSF_API int psf_scene_load_cache(psf_scene* scene, const char* path);

/* ---- structure tensor from any basis, on the device -------------------------------
 * Replaces the voxel loop of PSFSolver::buildAdvection (psfsolver.cpp:431-551)
 * for any basis (eigensolved, obstacle scenes); psf_scene_build_box_advection
 * covers the analytic closed box.  Inputs as buildAdvection holds them:
 * Phi (F x K, column-major FP64, velBasisField), V = curl Phi (nE x K,
 * column-major FP64, psfsolver.cpp:342), lambda[K] (finite, non-zero), and
 * one term per (inside voxel, edge slot) of psfsolver.cpp:456-533: its edge,
 * its faces a and b, and the coefficient s_h * c(a, b).  C_i = (1/lambda_i)
 * V^T W_i, W_i[e][j] = sum over e's terms of coef (Phi[a][i] Phi[b][j] -
 * Phi[a][j] Phi[b][i]), is formed on the GPU (FP64 matrix cores) and stored
 * in the scene's C precision (FP16 is refused: upload it).  Synchronous;
 * allocates Phi, V, the terms and a few GB of wedge matrices for the call. */
PSF_API int psf_scene_build_advection(psf_scene* scene, const double* Phi, const double* V, uint32_t nE,
                                      const double* lambda, uint32_t nterms, const uint32_t* term_edge,
                                      const uint32_t* term_a, const uint32_t* term_b, const double* term_coef);
/* C back to the host in the upload layout (K matrices C_i, column-major), for
 * scenes whose C precision is FP64 or FP32 (others store it packed:
 * PSF_ERR_UNSUPPORTED).  No counterpart in the reference; it lets a C built on
 * the device go into a scene cache (psf_cache_write).  Synchronous. */
PSF_API int psf_scene_get_advection(psf_scene* scene, double* C);

/* ---- eigensolve for the basis ---------------------------------------------------
 * Replaces the shift-invert eigensolve of the pruned Laplacian in
 * PSFSolver::buildLaplace (psfsolver.cpp:267-341, Spectra SymEigsShiftSolver
 * with one SparseLU per shift, eigenvalues above 1e-10 accepted).
 * The K smallest eigenpairs of a symmetric n x n A in CSR (row_ptr[n+1],
 * col_idx[nnz], values[nnz], host arrays), by Chebyshev-filtered subspace
 * iteration on the context's GPU: a block of K + guard Gaussian vectors,
 * filters of the given polynomial degree, Rayleigh-Ritz until every one of
 * the K residuals ||A x - lambda x|| is at most tol * (Gershgorin bound of
 * A's largest eigenvalue).  A null space is returned like any eigenspace;
 * lift an unwanted one with a penalty term first.
 * Out: lambda[K] ascending, X (n x K, column-major, orthonormal columns; the
 * sign and the basis inside a repeated eigenvalue are arbitrary), and the
 * number of products with A used.  Limits: K + guard <= 1024 and < n.  PSF_ERR_STATE if max_products is reached first.  Synchronous. */
PSF_API int psf_eigensolve_csr(psf_ctx* ctx, uint32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                               const double* values, uint32_t K, uint32_t guard, uint32_t degree, double tol,
                               uint32_t max_products, uint64_t seed, double* lambda, double* X,
                               uint32_t* products);

/* ---- particle density volume --------------------------------------------------
 * The histogram -> volume step of the GPU solver's particle drawing
 * (Solver/SolverImpl/psfsolvergpu.cpp:876-913, Resources/Volume/volume.comp:1-17;
 * 256 x 256 x 256 there).  Counts of particles per cell of an rx x ry x rz grid
 * spanning the scene's AABB, cell index (z * ry + y) * rx + x, and the volume
 * half(float(count) / 1e6) as IEEE binary16 bits (round to nearest even).
 * A particle's cell on each axis is trunc((p - aabb_min) / extent * r) in
 * float, clamped to [0, r-1] (NaN: 0), so every particle is counted once; the
 * reference's counting shader (history.comp) is absent from its sources, so
 * this rule is the library's own.
 * psf_set_density_grid allocates the grid (1..1024 cells per axis; 0,0,0
 * removes it).  With count_in_trace, every particle pass also counts the
 * particles' new positions, so psf_density after a frame reads no particle
 * records.  psf_density computes the density of the stored particles and
 * copies counts (rx*ry*rz uint32) and volume (rx*ry*rz uint16) to the host
 * where the pointers are not null; the volume stays on the device either way. */
PSF_API int psf_set_density_grid(psf_scene* scene, uint32_t rx, uint32_t ry, uint32_t rz, int count_in_trace);
PSF_API int psf_density(psf_scene* scene, uint32_t* counts, uint16_t* volume);

/* ---- measurement -----------------------------------------------------------
 * Runs `count` launches of one phase back to back on the context stream
 * between two HIP events and returns the elapsed milliseconds.  The timed
 * region is bracketed by two launches of an empty kernel,
 * k_timed_region_mark, so that kernel traces show what it contains.
 * phase: 0 reduced step (one step per launch), 1 reconstruction, 2 trace,
 *        3 `count` full integrate frames from a frame-window boundary: the
 *          frames computed ahead are dropped before the start event, window
 *          fills compute no frame past the count, and the end event follows
 *          every stream's work, so the region holds exactly the work of the
 *          `count` frames (their reduced steps, velocity fields and particle
 *          passes),
 *        4 one launch of `count` reduced steps,
 *        5 the frame window's reconstruction (all W frames; window > 1,
 *          recomputes the u of the frames already computed ahead),
 *        6 a density count pass of the stored particles (adds to the
 *          counters), 7 a density volume pass (clears the counters),
 *        8 a full integrate frame followed by psf_density without copies,
 *        9 one launch of the particle passes the windowed frame loop fuses
 *          (8 frames; needs a frame window of at least 8),
 *       10 `count` integrate() frames continuing the current frame window
 *          (as a caller's frame loop runs them; frames computed ahead before
 *          the start event are not in the region), then the deferred
 *          particle passes: count = 1 times one call's frame,
 *       11 phase 3 with every kernel launch of the region between two events
 *          of its own; psf_scene_region_kernels then reports their sums
 *          (the events lengthen the region, which phase 3 runs without). */
PSF_API int psf_time_phase(psf_scene* scene, int phase, int count, double* ms);

/* ---- scene server (continuous batching) ----------------------------------------
 * No counterpart in the reference, which runs one scene per solver.  A server
 * keeps a set of scenes on the device and advances all of them together: each
 * round runs up to `window` frames of every active scene -- one persistent
 * launch for all their reduced steps, one launch for all their velocity
 * fields, one launch for all their particle passes.  Requests wait in a FIFO
 * queue and are admitted at the start of a round when a slot, workgroups of
 * the persistent launch and device memory (the request's footprint against
 * mem_budget) are free; finished ones leave at the end of their round.
 * Scenes share one basis per grid (Phi, C of K0 modes); a request uses its
 * first K modes.  With shared_basis = 0 each request's C (K modes) is packed
 * on the device and held in registers by its own workgroups; with
 * shared_basis = 1 a basis's C is packed once (K0 <= 256) and one
 * register-resident copy serves up to 16 of its scenes, each a pair of B
 * columns of the same MFMA blocks.  A request's result does not depend on
 * what else runs in its rounds.  Frames are integrate() frames
 * (psfsolver.cpp:21-234) with C in BF16 hi+lo, FP32 Phi and FP64 omega. */
typedef struct psf_server_config {
    uint32_t window;        /* frames per round: 2, 4, 8, 16 or 32 (0 = 32) */
    uint32_t max_active;    /* scenes on the device at once, 1..64 (0 = 32) */
    uint64_t mem_budget;    /* device bytes for admitted requests (0 = 64 GiB) */
    uint32_t modes_per_wg;  /* modes per workgroup of the persistent layout, 1..8 (0 = per request, the
                               most that keep all of its C blocks in registers) */
    uint32_t u_fp32_sums;   /* 0: velocity fields u = Phi w with FP64 sums (PSF_PHI_FP32); 1: FP32 sums
                               on the FP32 matrix cores (PSF_PHI_FP32_ACC32) */
    uint32_t shared_basis;  /* 0: each request's reduced steps read its own packed C (modes_per_wg applies);
                               1: the scenes of a basis read one shared copy of its C (K0 <= 256) */
    uint32_t wait_limit_ns; /* bound of each wait of the reduced steps' exchange between workgroups, in ns
                               (0 = 2 s); a request whose launch gave up fails with PSF_ERR_HIP */
} psf_server_config;

typedef struct psf_server_stats {
    uint64_t submitted, completed, failed, rounds, scene_frames;
    uint32_t active, queued;
    uint64_t mem_in_use, mem_peak;   /* bytes held by admitted requests */
    double busy_ms;                  /* wall time inside rounds */
    double admit_ms, retire_ms;      /* of which: admitting requests, reading back finished ones */
} psf_server_stats;

PSF_API int psf_server_create(psf_ctx* ctx, const psf_server_config* cfg, psf_server** out);
PSF_API int psf_server_destroy(psf_server* server);
/* A shared basis: desc gives the grid and K0 = desc->K (c_precision and
   phi_precision are ignored); Phi F x K0 column-major, lambda and gravity K0,
   C as for psf_scene_upload_advection with K0 modes. */
PSF_API int psf_server_add_basis(psf_server* server, const psf_scene_desc* desc, const double* Phi,
                                 const double* lambda, const double* gravity, const double* C, uint32_t* basis_id);
/* Device bytes a request would hold (its admission footprint). */
PSF_API int psf_server_request_footprint(psf_server* server, uint32_t basis_id, uint32_t K, uint32_t n_particles,
                                         uint64_t* bytes);
/* Queue a request: nframes frames of the first K <= K0 modes of the basis,
   starting from omega (K) and n particles (x, y, z, w); the caller's buffers
   are copied. */
PSF_API int psf_server_submit(psf_server* server, uint32_t basis_id, uint32_t K, uint32_t nframes, double dt,
                              double viscosity, int gravity_active, float lifetime, const double* omega,
                              const float* particles, uint32_t n_particles, uint64_t* request_id);
/* One round: admit, advance every active scene, retire the finished ones.
   *completed receives the number of requests that finished in this round. */
PSF_API int psf_server_round(psf_server* server, uint32_t* completed);
/* 1 if the request finished (its result is ready), 0 if queued or running. */
PSF_API int psf_server_done(psf_server* server, uint64_t request_id);
/* The ids of requests that finished (done or failed) since the last call, in
   the order they finished: up to `capacity` of them into ids, their number
   into *count; the rest stay for the next call. */
PSF_API int psf_server_completed(psf_server* server, uint64_t* ids, uint32_t capacity, uint32_t* count);
/* Result of a finished request: omega (K), particles (n x 4, caller's order),
   latency from submit to the end of its last round in milliseconds.  The
   request is forgotten afterwards. */
PSF_API int psf_server_result(psf_server* server, uint64_t request_id, double* omega, float* particles,
                              double* latency_ms);
PSF_API int psf_server_get_stats(psf_server* server, psf_server_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* PSF_PSF_H */
