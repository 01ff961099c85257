// Host-side semantics of psf::AbstractSolver (include/psf/solver.hpp) against
// the reference's AbstractSolver (Solver/abstractsolver.cpp): particle ring,
// rebuild calls, benchmark.csv layout.  No GPU call is made.
//   solver_semantics <csv path>
// prints one "key value..." line per check.
#include <cstdio>

#include "psf/solver.hpp"

namespace {

class Probe : public psf::AbstractSolver {
public:
    int laplace = 0, eigen = 0, advection = 0, voxelized = 0;
    std::vector<std::pair<double, double>> frames;
    void configure(unsigned maxP, unsigned K, const char* csv)
    {
        maxParticles = maxP;
        nEigenFunctions = K;
        dims[0] = 4; dims[1] = 5; dims[2] = 6;
        for (int c = 0; c < 3; ++c) { aabbMin[c] = -0.5f; aabbMax[c] = 0.5f; }
        maxFramesBenchmark = 3;
        benchmarkPath = csv;
        clearParticles();
    }
    void integrate() override
    {
        benchmarkFrameBegin();
        benchmarkFrameTimes(1.5, 2.25);
        simTime += 1.0;
    }
    unsigned pointer() const { return particlePointer; }

protected:
    void buildLaplace() override { ++laplace; }
    void buildEigenFunctions() override { ++eigen; }
    void buildAdvection() override { ++advection; }
    void voxelize() override { ++voxelized; }
};

void dump(const char* key, const std::vector<float>& p)
{
    std::printf("%s", key);
    for (float v : p) std::printf(" %g", v);
    std::printf("\n");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 2) return 2;
    Probe s;
    s.configure(3, 8, argv[1]);
    dump("cleared", s.getParticles());
    s.addParticle(0.1f, 0.2f, 0.3f, 5.0f);     // slot 0: w = 0 <= simTime 0: taken, w = 0 + 5
    s.addParticle(9.0f, 0.0f, 0.0f, 1.0f);     // outside the AABB: ignored
    s.addParticle(0.0f, 0.1f, 0.0f, 2.0f);     // slot 1
    s.addParticle(0.0f, 0.0f, 0.1f, 3.0f);     // slot 2, pointer wraps to 0
    s.addParticle(-0.1f, 0.0f, 0.0f, 1.0f);    // slot 0 holds w = 5 > simTime 0: refused
    dump("added", s.getParticles());
    std::printf("pointer %u\n", s.pointer());
    for (int f = 0; f < 5; ++f) s.integrate();  // simTime 5
    s.addParticle(-0.1f, 0.0f, 0.0f, 1.0f);    // now 5 >= 5: slot 0 taken, w = 6
    dump("after5", s.getParticles());
    std::printf("simtime %g\n", s.getSimTime());
    s.setNumberEigenFunctions(12);
    std::printf("rebuild %d %d %d %d K %u\n", s.laplace, s.eigen, s.advection, s.voxelized, s.getNumEigenFunctions());
    s.setResolution(0.05);
    std::printf("rebuild %d %d %d %d simtime %g\n", s.laplace, s.eigen, s.advection, s.voxelized, s.getSimTime());
    s.changeNumParticles(2);
    std::printf("particles %u\n", s.getNumParticles());
    s.startBenchmark(true);
    for (int f = 0; f < 6; ++f) s.integrate();  // 3 frames recorded, the 5th frame ends it
    std::printf("benchmarking %d\n", s.benchmarking() ? 1 : 0);
    return 0;
}
