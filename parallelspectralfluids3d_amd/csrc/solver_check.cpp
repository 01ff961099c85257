// solver_check: drives psf::HipSolver (include/psf/solver.hpp) the way the
// reference's GLCanvas drives its solver: preprocessing fills the protected
// state (here read from files), buildAdvection() uploads it, integrate() runs
// once per frame, getters read the results back.
//
//   solver_check <dir> [extras]
// <dir>/meta.txt: nx ny nz resolution K N frames dt viscosity gravity lifetime c_precision phi_precision
// <dir>/{phi,lambda,gravity,C,omega}.bin (float64), parts.bin (float32 N x 4)
// writes <dir>/{omega_out.bin (float64), u_out.bin, parts_out.bin (float32)}
// With "extras" it then also writes the 16^3 density volume of the final
// particles (<dir>/volume_out.bin, uint16 half bits) and runs the reference's
// benchmark (startBenchmark, 4 recorded frames) into <dir>/benchmark.csv.
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

int main(int argc, char** argv)
{
    if (argc != 2 && argc != 3) {
        std::fprintf(stderr, "usage: %s <dir> [extras]\n", argv[0]);
        return 2;
    }
    const bool extras = argc == 3 && std::string(argv[2]) == "extras";
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
        for (unsigned f = 0; f < frames; ++f) solver.integrate();
        write_bin(dir + "/omega_out.bin", solver.getBasisCoefficients());
        write_bin(dir + "/u_out.bin", solver.getVelocityField());
        write_bin(dir + "/parts_out.bin", solver.getParticles());
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
