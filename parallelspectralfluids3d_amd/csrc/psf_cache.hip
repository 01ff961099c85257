// Scene cache: a versioned file holding a scene's precomputed inputs, the
// basis Phi, eigenvalues lambda, forcing projection g and structure tensor C
// (what buildLaplace/buildEigenFunctions/buildAdvection produce,
// psfsolver.cpp:236-551), so a scene is rebuilt by reading a file instead of
// an eigensolve and a C build.  Host code only; uploads go through the
// scene's ordinary upload calls, so a scene loaded from a cache is bitwise
// the scene those arrays would give.
//
// File layout (little-endian):
//   header  (CacheHeader below, 128 bytes)
//   lambda  K doubles
//   gravity K doubles                       if flags & kHasGravity
//   C       K^3 doubles, C[i][b][a] = C_i(a, b)
//   Phi     F * K doubles, column-major      if flags & kHasPhi
// The header's checksum is FNV-1a (64 bit) over every payload byte.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "psf_internal.hpp"

namespace psf {
namespace {

constexpr char kMagic[8] = {'P', 'S', 'F', 'C', 'A', 'C', 'H', 'E'};
constexpr uint32_t kFormat = 1;
constexpr uint32_t kHasPhi = 1u, kHasGravity = 2u;

struct CacheHeader {
    char magic[8];
    uint32_t format;            // kFormat
    uint32_t header_bytes;      // sizeof(CacheHeader)
    uint32_t nx, ny, nz, K;
    double resolution;
    float aabb_min[3], aabb_max[3];
    uint32_t flags;
    uint32_t reserved0;
    uint64_t F;                 // faces of the grid
    uint64_t payload_bytes;
    uint64_t checksum;          // FNV-1a over the payload
    uint8_t reserved[128 - 96];
};
static_assert(sizeof(CacheHeader) == 128, "cache header is 128 bytes");

uint64_t faces_of(uint32_t nx, uint32_t ny, uint32_t nz)
{
    return uint64_t(nx) * ny * (nz + 1) + uint64_t(nx) * (ny + 1) * nz + uint64_t(nx + 1) * ny * nz;
}

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n)
    {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    }
};

struct File {
    FILE* f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

int fail(psf_ctx* ctx, int code, const std::string& msg) { return ctx ? ctx_set_error(ctx, code, msg) : code; }

// reads and checks the header; the payload is not read
int read_header(FILE* f, CacheHeader& h, psf_ctx* ctx)
{
    if (std::fread(&h, sizeof(h), 1, f) != 1) return fail(ctx, PSF_ERR_ARG, "scene cache: short header");
    if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(ctx, PSF_ERR_ARG, "scene cache: not a scene cache file");
    if (h.format != kFormat || h.header_bytes != sizeof(CacheHeader))
        return fail(ctx, PSF_ERR_UNSUPPORTED, "scene cache: format version " + std::to_string(h.format) +
                                                  " (this library reads " + std::to_string(kFormat) + ")");
    if (h.K == 0 || h.nx == 0 || h.ny == 0 || h.nz == 0 || h.F != faces_of(h.nx, h.ny, h.nz))
        return fail(ctx, PSF_ERR_ARG, "scene cache: inconsistent header");
    const uint64_t K = h.K;
    const uint64_t expect = 8 * (K + ((h.flags & kHasGravity) ? K : 0) + K * K * K + ((h.flags & kHasPhi) ? h.F * K : 0));
    if (h.payload_bytes != expect) return fail(ctx, PSF_ERR_ARG, "scene cache: payload size does not match the header");
    return PSF_OK;
}

// reads n doubles into out, adding them to the checksum
int read_block(FILE* f, double* out, size_t n, Fnv& sum, psf_ctx* ctx)
{
    if (n && std::fread(out, sizeof(double), n, f) != n) return fail(ctx, PSF_ERR_ARG, "scene cache: file is truncated");
    sum.add(out, n * sizeof(double));
    return PSF_OK;
}

}  // namespace
}  // namespace psf

using namespace psf;

extern "C" {

int psf_cache_write(const char* path, const psf_scene_desc* desc, const double* Phi, const double* lambda,
                    const double* gravity, const double* C)
{
    if (!path || !desc || !lambda || !C) return PSF_ERR_ARG;
    if (desc->K == 0 || desc->nx == 0 || desc->ny == 0 || desc->nz == 0) return PSF_ERR_ARG;
    CacheHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.format = kFormat;
    h.header_bytes = sizeof(CacheHeader);
    h.nx = desc->nx;
    h.ny = desc->ny;
    h.nz = desc->nz;
    h.K = desc->K;
    h.resolution = desc->resolution;
    for (int a = 0; a < 3; ++a) {
        h.aabb_min[a] = desc->aabb_min[a];
        h.aabb_max[a] = desc->aabb_max[a];
    }
    h.flags = (Phi ? kHasPhi : 0u) | (gravity ? kHasGravity : 0u);
    h.F = faces_of(h.nx, h.ny, h.nz);
    const uint64_t K = h.K;
    const size_t nC = size_t(K * K * K), nPhi = Phi ? size_t(h.F * K) : 0;
    h.payload_bytes = 8 * (K + (gravity ? K : 0) + nC + nPhi);
    Fnv sum;
    sum.add(lambda, K * 8);
    if (gravity) sum.add(gravity, K * 8);
    sum.add(C, nC * 8);
    if (Phi) sum.add(Phi, nPhi * 8);
    h.checksum = sum.h;
    // write to a temporary name and rename, so a reader never sees a partial file
    const std::string tmp = std::string(path) + ".tmp";
    {
        File f;
        f.f = std::fopen(tmp.c_str(), "wb");
        if (!f.f) return PSF_ERR_ARG;
        bool ok = std::fwrite(&h, sizeof(h), 1, f.f) == 1 && std::fwrite(lambda, 8, K, f.f) == K;
        if (ok && gravity) ok = std::fwrite(gravity, 8, K, f.f) == K;
        if (ok) ok = std::fwrite(C, 8, nC, f.f) == nC;
        if (ok && Phi) ok = std::fwrite(Phi, 8, nPhi, f.f) == nPhi;
        if (ok) ok = std::fflush(f.f) == 0;
        if (!ok) {
            std::fclose(f.f);
            f.f = nullptr;
            std::remove(tmp.c_str());
            return PSF_ERR_ARG;
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return PSF_ERR_ARG;
    }
    return PSF_OK;
}

int psf_cache_read_header(const char* path, psf_scene_desc* desc, uint32_t* has_phi, uint32_t* has_gravity)
{
    if (!path || !desc) return PSF_ERR_ARG;
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return PSF_ERR_ARG;
    CacheHeader h;
    if (int rc = read_header(f.f, h, nullptr)) return rc;
    desc->nx = h.nx;
    desc->ny = h.ny;
    desc->nz = h.nz;
    desc->K = h.K;
    desc->resolution = h.resolution;
    for (int a = 0; a < 3; ++a) {
        desc->aabb_min[a] = h.aabb_min[a];
        desc->aabb_max[a] = h.aabb_max[a];
    }
    if (has_phi) *has_phi = (h.flags & kHasPhi) ? 1u : 0u;
    if (has_gravity) *has_gravity = (h.flags & kHasGravity) ? 1u : 0u;
    return PSF_OK;
}

int psf_cache_read(const char* path, double* Phi, double* lambda, double* gravity, double* C)
{
    if (!path || !lambda || !C) return PSF_ERR_ARG;
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return PSF_ERR_ARG;
    CacheHeader h;
    if (int rc = read_header(f.f, h, nullptr)) return rc;
    const size_t K = h.K;
    std::vector<double> scratch;
    Fnv sum;
    if (int rc = read_block(f.f, lambda, K, sum, nullptr)) return rc;
    if (h.flags & kHasGravity) {
        double* g = gravity;
        if (!g) {
            scratch.resize(K);
            g = scratch.data();
        }
        if (int rc = read_block(f.f, g, K, sum, nullptr)) return rc;
    }
    if (int rc = read_block(f.f, C, K * K * K, sum, nullptr)) return rc;
    if (h.flags & kHasPhi) {
        double* p = Phi;
        std::vector<double> skip;
        if (!p) {
            skip.resize(size_t(h.F) * K);
            p = skip.data();
        }
        if (int rc = read_block(f.f, p, size_t(h.F) * K, sum, nullptr)) return rc;
    }
    return sum.h == h.checksum ? PSF_OK : PSF_ERR_ARG;
}

int psf_scene_load_cache(psf_scene* s, const char* path)
{
    if (!s || !path) return PSF_ERR_ARG;
    psf_ctx* ctx = s->ctx;
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return fail(ctx, PSF_ERR_ARG, std::string("scene cache: cannot open ") + path);
    CacheHeader h;
    if (int rc = read_header(f.f, h, ctx)) return rc;
    if (h.nx != s->d.nx || h.ny != s->d.ny || h.nz != s->d.nz || h.K != s->d.K || h.resolution != s->d.resolution)
        return fail(ctx, PSF_ERR_ARG, "scene cache: grid, resolution or K differ from the scene's");
    const size_t K = h.K;
    Fnv sum;
    std::vector<double> lambda(K), gravity, C(K * K * K), Phi;
    if (int rc = read_block(f.f, lambda.data(), K, sum, ctx)) return rc;
    if (h.flags & kHasGravity) {
        gravity.resize(K);
        if (int rc = read_block(f.f, gravity.data(), K, sum, ctx)) return rc;
    }
    if (int rc = read_block(f.f, C.data(), C.size(), sum, ctx)) return rc;
    if (h.flags & kHasPhi) {
        Phi.resize(size_t(h.F) * K);
        if (int rc = read_block(f.f, Phi.data(), Phi.size(), sum, ctx)) return rc;
    }
    if (sum.h != h.checksum) return fail(ctx, PSF_ERR_ARG, "scene cache: checksum mismatch (corrupt file)");
    if (int rc = psf_scene_upload_basis(s, Phi.empty() ? nullptr : Phi.data(), lambda.data(),
                                        gravity.empty() ? nullptr : gravity.data()))
        return rc;
    return psf_scene_upload_advection(s, C.data());
}

}  // extern "C"
