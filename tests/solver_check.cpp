// solver_check: test driver (tests/test_gpu_solver_cpp.py) of psf::HipSolver
// (include/psf/solver.hpp), driven the way the reference's GLCanvas drives its
// solver: preprocessing fills the protected state (here read from files),
// buildAdvection() uploads it, integrate() runs once per frame, getters read
// the results back.
//
//   solver_check <dir> [extras | window <W>] | eigen | leak
// <dir>/meta.txt: nx ny nz resolution K N frames dt viscosity gravity lifetime c_precision phi_precision
// <dir>/{phi,lambda,gravity,C,omega}.bin (float64), parts.bin (float32 N x 4)
// writes <dir>/{omega_out.bin (float64), u_out.bin, parts_out.bin (float32)}
// With "extras" it then also writes the 16^3 density volume of the final
// particles (<dir>/volume_out.bin, uint16 half bits) and runs the reference's
// benchmark (startBenchmark, 4 recorded frames) into <dir>/benchmark.csv.
// With "window W" the solver runs with a frame window of W and, like the
// reference's canvas loop, sets the (unchanged) timestep before every
// integrate(); it prints the scene's counters.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "psf/solver.hpp"

namespace {

template <typename T>
std::vector<T> read_bin(const std::string& path, size_t n)
{
    std::vector<T> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T))))
        throw std::runtime_error("cannot read " + path);
    return v;
}

template <typename T>
void write_bin(const std::string& path, const std::vector<T>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

class FileSolver : public psf::HipSolver {
public:
    FileSolver(const std::string& dir, int cprec, int phiprec) : psf::HipSolver(0, cprec, phiprec), dir_(dir) {}

    void configure(uint32_t nx, uint32_t ny, uint32_t nz, double res, unsigned K, unsigned N)
    {
        dims[0] = nx; dims[1] = ny; dims[2] = nz;
        resolution = res;
        nEigenFunctions = K;
        maxParticles = N;
        for (int c = 0; c < 3; ++c) {
            const float half = static_cast<float>(0.5 * dims[c] * res);
            aabbMin[c] = -half;
            aabbMax[c] = half;
        }
        buildLaplace();
        buildAdvection();
    }

    void benchmarkSetup(const std::string& path, unsigned frames)
    {
        benchmarkPath = path;
        maxFramesBenchmark = frames;
    }

protected:
    void buildLaplace() override
    {
        const size_t F = size_t(dims[0] + 1) * dims[1] * dims[2] + size_t(dims[1] + 1) * dims[0] * dims[2] +
                         size_t(dims[2] + 1) * dims[0] * dims[1];
        const size_t K = nEigenFunctions;
        velBasisField = read_bin<double>(dir_ + "/phi.bin", F * K);
        eigenValues = read_bin<double>(dir_ + "/lambda.bin", K);
        gravity = read_bin<double>(dir_ + "/gravity.bin", K);
        basisCoeff = read_bin<double>(dir_ + "/omega.bin", K);
        particles = read_bin<float>(dir_ + "/parts.bin", size_t(maxParticles) * 4);
    }
    void buildAdvection() override
    {
        const size_t K = nEigenFunctions;
        advection = read_bin<double>(dir_ + "/C.bin", K * K * K);
        upload();
    }

private:
    std::string dir_;
};

}  // namespace

// psf_eigensolve_csr from C++: the Dirichlet path Laplacian (tridiagonal
// 2, -1), whose eigenpairs are 2 - 2 cos(pi k / (n + 1)) and sine vectors
int eigen_check()
{
    const uint32_t n = 3000, K = 16;
    std::vector<int32_t> rp(n + 1), ci;
    std::vector<double> av;
    for (uint32_t i = 0; i < n; ++i) {
        rp[i] = static_cast<int32_t>(ci.size());
        if (i > 0) { ci.push_back(static_cast<int32_t>(i - 1)); av.push_back(-1.0); }
        ci.push_back(static_cast<int32_t>(i)); av.push_back(2.0);
        if (i + 1 < n) { ci.push_back(static_cast<int32_t>(i + 1)); av.push_back(-1.0); }
    }
    rp[n] = static_cast<int32_t>(ci.size());
    psf_ctx* ctx = nullptr;
    if (psf_ctx_create(0, &ctx) != PSF_OK) {
        std::fprintf(stderr, "psf_ctx_create failed\n");
        return 1;
    }
    std::vector<double> lam(K), X(size_t(n) * K);
    uint32_t products = 0;
    const int rc = psf_eigensolve_csr(ctx, n, rp.data(), ci.data(), av.data(), K, 16, 30, 1e-12, 20000, 7, lam.data(),
                                      X.data(), &products);
    if (rc != PSF_OK) {
        std::fprintf(stderr, "psf_eigensolve_csr: %d %s\n", rc, psf_last_error(ctx));
        psf_ctx_destroy(ctx);
        return 1;
    }
    double err = 0.0, vec = 0.0;
    const double pi = 3.14159265358979323846;
    for (uint32_t k = 0; k < K; ++k) {
        err = std::max(err, std::fabs(lam[k] - (2.0 - 2.0 * std::cos(pi * (k + 1) / (n + 1)))));
        double dot = 0.0, nrm = 0.0;
        for (uint32_t j = 0; j < n; ++j) {
            const double v = std::sin(pi * (k + 1) * (j + 1) / (n + 1));
            dot += v * X[size_t(k) * n + j];
            nrm += v * v;
        }
        vec = std::max(vec, std::fabs(std::fabs(dot) / std::sqrt(nrm) - 1.0));
    }
    psf_ctx_destroy(ctx);
    std::printf("eigen_check: %u products, max eigenvalue error %.3e, max |cos| - 1 %.3e\n", products, err, vec);
    return (err < 1e-12 && vec < 1e-9) ? 0 : 1;
}

// A context and a scene with queued work left alive at exit: the library
// releases them itself, before the HIP runtime's teardown, and the process
// exits normally.
int leak_check()
{
    psf_ctx* ctx = nullptr;
    if (psf_ctx_create(0, &ctx) != PSF_OK) return 1;
    psf_scene_desc d{};
    d.nx = d.ny = d.nz = 8;
    d.resolution = 0.125;
    for (int c = 0; c < 3; ++c) {
        d.aabb_min[c] = -0.5f;
        d.aabb_max[c] = 0.5f;
    }
    d.K = 4;
    d.c_precision = PSF_C_BF16X2;
    d.phi_precision = PSF_PHI_FP32;
    d.max_particles = 0;
    psf_scene* s = nullptr;
    if (psf_scene_create(ctx, &d, &s) != PSF_OK) return 1;
    std::vector<double> lam = {1.0, 2.0, 3.0, 4.0}, C(64, 0.0), w = {0.5, 0.5, 0.5, 0.5};
    if (psf_scene_upload_basis(s, nullptr, lam.data(), nullptr) != PSF_OK) return 1;
    if (psf_scene_upload_advection(s, C.data()) != PSF_OK) return 1;
    if (psf_set_omega(s, w.data()) != PSF_OK) return 1;
    if (psf_step(s, 1000) != PSF_OK) return 1;   // still running when main returns
    std::printf("leak_check: returning with a live context\n");
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 2 && std::string(argv[1]) == "eigen") return eigen_check();
    if (argc == 2 && std::string(argv[1]) == "leak") return leak_check();
    if (argc != 2 && argc != 3 && !(argc == 4 && std::string(argv[2]) == "window")) {
        std::fprintf(stderr, "usage: %s <dir> [extras | window <W>] | eigen\n", argv[0]);
        return 2;
    }
    const bool extras = argc == 3 && std::string(argv[2]) == "extras";
    const unsigned window = argc == 4 ? static_cast<unsigned>(std::stoul(argv[3])) : 1u;
    const std::string dir = argv[1];
    std::ifstream meta(dir + "/meta.txt");
    uint32_t nx, ny, nz;
    double res, dt, nu;
    unsigned K, N, frames;
    int grav, cprec, phiprec;
    float lifetime;
    if (!(meta >> nx >> ny >> nz >> res >> K >> N >> frames >> dt >> nu >> grav >> lifetime >> cprec >> phiprec)) {
        std::fprintf(stderr, "bad meta.txt\n");
        return 2;
    }
    try {
        FileSolver solver(dir, cprec, phiprec);
        solver.configure(nx, ny, nz, res, K, N);
        solver.setTimestep(dt);
        solver.setViscosity(nu);
        solver.setGravityActive(grav != 0);
        solver.setLifeTime(lifetime / 60.0f);
        if (window > 1) solver.setFrameWindow(window);
        for (unsigned f = 0; f < frames; ++f) {
            if (window > 1) solver.setTimestep(dt);   // the canvas loop's per-frame parameters
            solver.integrate();
        }
        write_bin(dir + "/omega_out.bin", solver.getBasisCoefficients());
        write_bin(dir + "/u_out.bin", solver.getVelocityField());
        write_bin(dir + "/parts_out.bin", solver.getParticles());
        if (window > 1) {
            const psf_scene_counters c = solver.counters();
            std::printf("counters frames=%llu window_fills=%llu window_frames=%llu windows_dropped=%llu "
                        "step_launches=%llu recon_launches=%llu particle_launches=%llu\n",
                        (unsigned long long)c.frames, (unsigned long long)c.window_fills,
                        (unsigned long long)c.window_frames, (unsigned long long)c.windows_dropped,
                        (unsigned long long)c.step_launches, (unsigned long long)c.recon_launches,
                        (unsigned long long)c.particle_launches);
        }
        if (extras) {
            solver.enableDensity(16, 16, 16);
            write_bin(dir + "/volume_out.bin", solver.getDensityVolume());
            solver.benchmarkSetup(dir + "/benchmark.csv", 3);
            solver.startBenchmark(true);
            for (int f = 0; f < 5; ++f) solver.integrate();
            if (solver.benchmarking()) throw std::runtime_error("benchmark did not end");
        }
    } catch (const psf::Error& e) {
        std::fprintf(stderr, "psf error %d: %s\n", e.status(), e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("solver_check: %u frames done\n", frames);
    return 0;
}
