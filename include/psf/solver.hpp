// solver.hpp -- C++ mirror of the reference's solver interface over the C ABI.
//
// psf::AbstractSolver restates the part of AbstractSolver
// (Solver/abstractsolver.h:13-127) that the per-frame step reads: the
// parameters set through setTimestep/setViscosity/setGravityActive/setLifeTime
// (abstractsolver.cpp:226-250), and the protected members that buildLaplace()
// and buildAdvection() fill for integrate() (abstractsolver.h:94-126), minus
// OpenGL, Qt and Eigen.  psf::HipSolver is the SolverImpl that replaces
// PSFSolverGPU (Solver/SolverImpl/psfsolvergpu.h:10-88): a subclass fills
// the state in buildLaplace()/buildAdvection() as the reference does, calls
// upload(), and integrate() then runs on the GPU.
//
// Layouts are Eigen's: velBasisField F x K column-major, advection K
// matrices K x K column-major (matrix i at advection.data() + i*K*K),
// particles as (x, y, z, w) records.  Errors throw psf::Error carrying the
// psf_status code and the library's message; nothing exits.
//
// The rebuild and particle semantics of the reference are kept:
// setNumberEigenFunctions / setResolution clear the particles and rerun the
// preprocessing (abstractsolver.cpp:178-222), changeNumParticles resizes the
// particle ring (:241-245), clearParticles zeroes it (:379-385) and
// addParticle fills it round-robin, taking a slot only once its record's w
// (an expiry frame) has passed (:387-397).  startBenchmark writes the
// reference's benchmark.csv (:252-302): particle count, K, grid, the three
// preprocessing times, then one "reduced step + reconstruction, particle pass"
// line in milliseconds per frame for maxFramesBenchmark frames, then their
// averages.
#pragma once

#include <chrono>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "psf/psf.h"

namespace psf {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

class AbstractSolver {
public:
    virtual ~AbstractSolver() = default;

    virtual void integrate() = 0;                                   // abstractsolver.h:18

    void setTimestep(double timestep) { timeStep = timestep; paramsChanged(); }
    void setViscosity(double visc) { viscosity = visc; paramsChanged(); }
    void setGravityActive(bool state) { gravityActive = state; paramsChanged(); }
    void setLifeTime(float lt) { lifeTime = lt * 60.0f; paramsChanged(); }   // abstractsolver.cpp:246-249

    unsigned int getNumEigenFunctions() const { return nEigenFunctions; }
    double getTimestep() const { return timeStep; }
    virtual const std::vector<double>& getBasisCoefficients() { return basisCoeff; }
    virtual const std::vector<float>& getVelocityField() { return velocityField; }
    virtual const std::vector<float>& getParticles() { return particles; }
    unsigned int getNumParticles() const { return static_cast<unsigned int>(particles.size() / 4); }
    double getSimTime() const { return simTime; }

    // abstractsolver.cpp:178-185
    void setNumberEigenFunctions(unsigned int n)
    {
        clearParticles();
        nEigenFunctions = n;
        rebuild();
    }
    // abstractsolver.cpp:187-222: voxelize() stands for Model::voxelize
    void setResolution(double res)
    {
        clearParticles();
        resolution = res;
        voxelize();
        rebuild();
    }
    // abstractsolver.cpp:241-245
    void changeNumParticles(unsigned int n)
    {
        maxParticles = n;
        clearParticles();
    }
    // abstractsolver.cpp:379-385: maxParticles zero records
    void clearParticles()
    {
        simTime = 0.0;
        particlePointer = 0;
        particles.assign(size_t(maxParticles) * 4, 0.0f);
        particlesWritten();
    }
    // abstractsolver.cpp:387-397: the slot at particlePointer takes the
    // particle once simTime has reached its record's w; the new record's w is
    // simTime + the particle's w
    void addParticle(float x, float y, float z, float w)
    {
        if (maxParticles == 0) return;
        particlesRead();
        float* slot = &particles[size_t(particlePointer) * 4];
        if (isPointInside(x, y, z) && simTime >= static_cast<double>(slot[3])) {
            slot[0] = x;
            slot[1] = y;
            slot[2] = z;
            slot[3] = static_cast<float>(simTime + static_cast<double>(w));
            particlePointer = (particlePointer + 1) % maxParticles;
            particlesWritten();
        }
    }
    // abstractsolver.cpp:252-302
    void startBenchmark(bool benchmark)
    {
        using Clock = std::chrono::high_resolution_clock;
        auto ms = [](Clock::time_point a, Clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        if (benchmark) {
            isBenchmark = true;
            benchmarkSums.assign(2, 0.0);
            benchmarkFrameNo = 0;
            if (benchmarkFile.is_open()) benchmarkFile.close();
            benchmarkFile.open(benchmarkPath);
            benchmarkFile.setf(std::ios::fixed, std::ios::floatfield);
            benchmarkFile << maxParticles << std::endl;
            benchmarkFile << nEigenFunctions << std::endl;
            benchmarkFile << dims[0] << "," << dims[1] << "," << dims[2] << std::endl;
            clearParticles();
            Clock::time_point t0 = Clock::now();
            buildLaplace();
            Clock::time_point t1 = Clock::now();
            benchmarkFile << ms(t0, t1) << ",";
            buildEigenFunctions();
            Clock::time_point t2 = Clock::now();
            benchmarkFile << ms(t1, t2) << ",";
            buildAdvection();
            Clock::time_point t3 = Clock::now();
            benchmarkFile << ms(t2, t3) << std::endl << std::endl;
        } else {
            benchmarkFile << std::endl << std::endl;
            benchmarkFile << benchmarkSums[0] / maxFramesBenchmark << "," << benchmarkSums[1] / maxFramesBenchmark;
            isBenchmark = false;
            benchmarkFrameNo = 0;
            if (benchmarkFile.is_open()) benchmarkFile.close();
        }
    }
    bool benchmarking() const { return isBenchmark; }

protected:
    virtual void buildLaplace() = 0;                                 // abstractsolver.h:77
    virtual void buildEigenFunctions() {}                            // abstractsolver.h:79
    virtual void buildAdvection() = 0;                               // abstractsolver.h:78
    virtual void paramsChanged() {}
    // grid of the model at `resolution` (Model::voxelize): dims and the AABB
    virtual void voxelize() {}
    // DECMesh3D::isPointInside; the default is the AABB
    virtual bool isPointInside(float x, float y, float z) const
    {
        const float p[3] = {x, y, z};
        for (int c = 0; c < 3; ++c)
            if (!(p[c] >= aabbMin[c] && p[c] <= aabbMax[c])) return false;
        return true;
    }
    // hooks around host access to `particles` (a device-resident solver syncs
    // before reads and uploads after writes)
    virtual void particlesRead() {}
    virtual void particlesWritten() {}

    // the start of a frame under benchmark (psfsolver.cpp:23-30)
    void benchmarkFrameBegin()
    {
        if (isBenchmark && benchmarkFrameNo > maxFramesBenchmark) startBenchmark(false);
        if (isBenchmark) benchmarkFrameNo++;
    }
    // one frame's two times (psfsolver.cpp:62-67, 226-231)
    void benchmarkFrameTimes(double reducedMs, double particlesMs)
    {
        if (!isBenchmark) return;
        benchmarkSums[0] += reducedMs;
        benchmarkSums[1] += particlesMs;
        benchmarkFile << reducedMs << "," << particlesMs << std::endl;
    }

    void rebuild()
    {
        buildLaplace();
        buildEigenFunctions();
        buildAdvection();
    }

    // defaults of AbstractSolver::AbstractSolver (abstractsolver.cpp:7-33)
    unsigned int maxParticles = 1000;
    unsigned int nEigenFunctions = 16;
    double resolution = 0.1;
    double timeStep = 1.0 / 60.0;
    double viscosity = 0.0;
    bool gravityActive = false;
    float lifeTime = 60.0f;

    // frames since clearParticles (psfsolver.cpp:233) and the ring position
    double simTime = 0.0;
    unsigned int particlePointer = 0;

    // benchmark state (abstractsolver.h:66-71; maxFramesBenchmark from
    // abstractsolver.cpp:10)
    unsigned int maxFramesBenchmark = 1000;
    bool isBenchmark = false;
    unsigned int benchmarkFrameNo = 0;
    std::vector<double> benchmarkSums;
    std::ofstream benchmarkFile;
    std::string benchmarkPath = "benchmark.csv";

    // grid (DECMesh3D::getDimensions) and Model::getAABB()
    uint32_t dims[3] = {0, 0, 0};
    float aabbMin[3] = {0.f, 0.f, 0.f};
    float aabbMax[3] = {0.f, 0.f, 0.f};

    // the data contract between preprocessing and integrate()
    std::vector<double> basisCoeff;     // omega, K
    std::vector<double> eigenValues;    // lambda, K
    std::vector<double> gravity;        // Phi^T f, K
    std::vector<double> velBasisField;  // Phi, F x K column-major
    std::vector<double> advection;      // C, K x (K x K column-major)
    std::vector<float> velocityField;   // u, F
    std::vector<float> particles;       // N x (x, y, z, w)
};

class HipSolver : public AbstractSolver {
public:
    explicit HipSolver(int device = 0, int c_precision = PSF_C_BF16X2, int phi_precision = PSF_PHI_FP32)
        : cPrecision_(c_precision), phiPrecision_(phi_precision)
    {
        check(psf_ctx_create(device, &ctx_), "psf_ctx_create");
    }
    ~HipSolver() override
    {
        if (scene_) psf_scene_destroy(scene_);
        if (ctx_) psf_ctx_destroy(ctx_);
    }
    HipSolver(const HipSolver&) = delete;
    HipSolver& operator=(const HipSolver&) = delete;

    // The upload tail of PSFSolverGPU::buildAdvection (psfsolvergpu.cpp:718-874):
    // copies the state the preprocessing left in the protected members to the
    // device, where it stays; call it at the end of buildAdvection().
    void upload()
    {
        if (scene_) {
            psf_scene_destroy(scene_);
            scene_ = nullptr;
        }
        psf_scene_desc d{};
        d.nx = dims[0];
        d.ny = dims[1];
        d.nz = dims[2];
        d.resolution = resolution;
        for (int c = 0; c < 3; ++c) {
            d.aabb_min[c] = aabbMin[c];
            d.aabb_max[c] = aabbMax[c];
        }
        d.K = nEigenFunctions;
        d.c_precision = cPrecision_;
        d.phi_precision = phiPrecision_;
        if (particles.size() != size_t(maxParticles) * 4) particles.resize(size_t(maxParticles) * 4, 0.0f);
        d.max_particles = maxParticles;
        check(psf_scene_create(ctx_, &d, &scene_), "psf_scene_create");
        check(psf_scene_upload_basis(scene_, velBasisField.data(), eigenValues.data(),
                                     gravity.empty() ? nullptr : gravity.data()), "upload_basis");
        check(psf_scene_upload_advection(scene_, advection.data()), "upload_advection");
        check(psf_set_omega(scene_, basisCoeff.data()), "set_omega");
        check(psf_set_particles(scene_, particles.data(), d.max_particles), "set_particles");
        if (densityDims_[0]) check(psf_set_density_grid(scene_, densityDims_[0], densityDims_[1], densityDims_[2], 1),
                                   "set_density_grid");
        if (frameWindow_ > 1) check(psf_set_frame_window(scene_, frameWindow_), "set_frame_window");
        paramsChanged();
        deviceAhead_ = false;
        particlesDirty_ = false;
    }

    // One frame: reduced step, u = Phi w, particle pass (psfsolver.cpp:21-234).
    // Under benchmark the two parts run as separately timed launches.
    void integrate() override
    {
        if (!scene_) throw Error(PSF_ERR_STATE, "integrate() before upload()");
        benchmarkFrameBegin();
        if (particlesDirty_) {
            check(psf_set_particles(scene_, particles.data(), static_cast<uint32_t>(particles.size() / 4)),
                  "set_particles");
            particlesDirty_ = false;
        }
        if (isBenchmark) {
            double step = 0.0, rec = 0.0, tr = 0.0;
            check(psf_time_phase(scene_, 0, 1, &step), "time reduced step");
            check(psf_time_phase(scene_, 1, 1, &rec), "time reconstruction");
            check(psf_time_phase(scene_, 2, 1, &tr), "time particle pass");
            benchmarkFrameTimes(step + rec, tr);
        } else {
            check(psf_integrate(scene_, 1), "psf_integrate");
        }
        simTime += 1.0;
        deviceAhead_ = true;
    }

    // Frame window (psf_set_frame_window; no counterpart in the reference):
    // integrate() then computes W frames' reduced steps and velocity fields
    // at once, ahead of the particle passes, with the bits of single frames.
    // The reference's call pattern -- setTimestep and friends with unchanged
    // values, then integrate(), once per frame -- keeps the window.  Kept
    // across upload(); W is 1 (off), 2, 4, 8, 16 or 32.
    void setFrameWindow(uint32_t frames)
    {
        if (scene_) check(psf_set_frame_window(scene_, frames), "set_frame_window");
        else if (frames != 1 && frames != 2 && frames != 4 && frames != 8 && frames != 16 && frames != 32)
            throw Error(PSF_ERR_ARG, "frame window must be 1, 2, 4, 8, 16 or 32");
        frameWindow_ = frames;
    }
    uint32_t getFrameWindow() const { return frameWindow_; }
    // what the device path has queued (psf_scene_get_counters)
    psf_scene_counters counters()
    {
        psf_scene_counters c{};
        if (scene_) check(psf_scene_get_counters(scene_, &c), "get_counters");
        return c;
    }

    // The particle density volume the reference draws (psfsolvergpu.cpp:876-913):
    // enable once (256^3 there); getDensityVolume() returns half-precision
    // bits, x fastest.
    void enableDensity(uint32_t nx = 256, uint32_t ny = 256, uint32_t nz = 256)
    {
        densityDims_[0] = nx;
        densityDims_[1] = ny;
        densityDims_[2] = nz;
        if (scene_) check(psf_set_density_grid(scene_, nx, ny, nz, 1), "set_density_grid");
    }
    const std::vector<uint16_t>& getDensityVolume()
    {
        if (!scene_ || !densityDims_[0]) throw Error(PSF_ERR_STATE, "no density grid");
        if (particlesDirty_) {
            check(psf_set_particles(scene_, particles.data(), static_cast<uint32_t>(particles.size() / 4)),
                  "set_particles");
            particlesDirty_ = false;
        }
        volume_.resize(size_t(densityDims_[0]) * densityDims_[1] * densityDims_[2]);
        check(psf_density(scene_, nullptr, volume_.data()), "density");
        return volume_;
    }

    // Reduced steps only, n per launch (the persistent path keeps C on chip).
    void steps(int n)
    {
        if (!scene_) throw Error(PSF_ERR_STATE, "steps() before upload()");
        check(psf_step(scene_, n), "psf_step");
        deviceAhead_ = true;
    }

    const std::vector<double>& getBasisCoefficients() override
    {
        sync();
        return basisCoeff;
    }
    const std::vector<float>& getVelocityField() override
    {
        sync();
        return velocityField;
    }
    const std::vector<float>& getParticles() override
    {
        sync();
        return particles;
    }

protected:
    void paramsChanged() override
    {
        if (scene_) check(psf_set_params(scene_, timeStep, viscosity, gravityActive ? 1 : 0, lifeTime), "set_params");
    }
    void particlesRead() override { sync(); }
    void particlesWritten() override { particlesDirty_ = scene_ != nullptr; }

private:
    void check(int rc, const char* what)
    {
        if (rc != PSF_OK) throw Error(rc, std::string(what) + ": " + (ctx_ ? psf_last_error(ctx_) : "no context"));
    }
    // device results to the host members; particle records written on the
    // host since the last upload are newer than the device's and stay
    void sync()
    {
        if (!scene_ || !deviceAhead_) return;
        basisCoeff.resize(nEigenFunctions);
        check(psf_get_omega(scene_, basisCoeff.data()), "get_omega");
        velocityField.resize(size_t(dims[0] + 1) * dims[1] * dims[2] + size_t(dims[1] + 1) * dims[0] * dims[2] +
                             size_t(dims[2] + 1) * dims[0] * dims[1]);
        check(psf_get_velocity(scene_, velocityField.data()), "get_velocity");
        if (!particlesDirty_)
            check(psf_get_particles(scene_, particles.data(), static_cast<uint32_t>(particles.size() / 4)),
                  "get_particles");
        deviceAhead_ = false;
    }

    psf_ctx* ctx_ = nullptr;
    psf_scene* scene_ = nullptr;
    int cPrecision_, phiPrecision_;
    bool deviceAhead_ = false;
    bool particlesDirty_ = false;    // the host records changed since the last upload
    uint32_t densityDims_[3] = {0, 0, 0};
    uint32_t frameWindow_ = 1;
    std::vector<uint16_t> volume_;
};

}  // namespace psf
