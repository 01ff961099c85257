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
#pragma once

#include <cstdint>
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

protected:
    virtual void buildLaplace() = 0;                                 // abstractsolver.h:77
    virtual void buildAdvection() = 0;                               // abstractsolver.h:78
    virtual void paramsChanged() {}

    // defaults of AbstractSolver::AbstractSolver (abstractsolver.cpp:7-33)
    unsigned int maxParticles = 1000;
    unsigned int nEigenFunctions = 16;
    double resolution = 0.1;
    double timeStep = 1.0 / 60.0;
    double viscosity = 0.0;
    bool gravityActive = false;
    float lifeTime = 60.0f;

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
        d.max_particles = static_cast<uint32_t>(particles.size() / 4);
        check(psf_scene_create(ctx_, &d, &scene_), "psf_scene_create");
        check(psf_scene_upload_basis(scene_, velBasisField.data(), eigenValues.data(),
                                     gravity.empty() ? nullptr : gravity.data()), "upload_basis");
        check(psf_scene_upload_advection(scene_, advection.data()), "upload_advection");
        check(psf_set_omega(scene_, basisCoeff.data()), "set_omega");
        check(psf_set_particles(scene_, particles.data(), d.max_particles), "set_particles");
        paramsChanged();
        deviceAhead_ = false;
    }

    // One frame: reduced step, u = Phi w, particle pass (psfsolver.cpp:21-234).
    void integrate() override
    {
        if (!scene_) throw Error(PSF_ERR_STATE, "integrate() before upload()");
        check(psf_integrate(scene_, 1), "psf_integrate");
        deviceAhead_ = true;
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

private:
    void check(int rc, const char* what)
    {
        if (rc != PSF_OK) throw Error(rc, std::string(what) + ": " + (ctx_ ? psf_last_error(ctx_) : "no context"));
    }
    void sync()
    {
        if (!scene_ || !deviceAhead_) return;
        basisCoeff.resize(nEigenFunctions);
        check(psf_get_omega(scene_, basisCoeff.data()), "get_omega");
        velocityField.resize(size_t(dims[0] + 1) * dims[1] * dims[2] + size_t(dims[1] + 1) * dims[0] * dims[2] +
                             size_t(dims[2] + 1) * dims[0] * dims[1]);
        check(psf_get_velocity(scene_, velocityField.data()), "get_velocity");
        check(psf_get_particles(scene_, particles.data(), static_cast<uint32_t>(particles.size() / 4)), "get_particles");
        deviceAhead_ = false;
    }

    psf_ctx* ctx_ = nullptr;
    psf_scene* scene_ = nullptr;
    int cPrecision_, phiPrecision_;
    bool deviceAhead_ = false;
};

}  // namespace psf
