// Particle density volume (C ABI in psf.h).
//
// The GPU solver of the reference draws its particles as a volume
// (psfsolvergpu.cpp:876-913): it clears a 256^3 r32i image, runs a compute
// pass that counts the particles per cell of the AABB, and converts the
// counts to an r16f image, count / 1e6 (Resources/Volume/volume.comp:1-17).
// Here the counters and the volume are scene buffers of any grid size:
//   - a particle pass can count each particle's new position as it stores it
//     (no second read of the particle records), or a count pass reads the
//     stored particles;
//   - the volume pass converts the counters and clears them in the same
//     sweep, so the next count starts from zero without a fill.
// The counting shader (history.comp) is not part of the reference, so the
// cell rule is the library's: per axis trunc((p - aabb_min) / extent * n) in
// float, clamped to [0, n-1].
#include <string>

#include "psf_internal.hpp"
#include "psf_kernels.hpp"

namespace psf {
namespace {

int fail(psf_ctx* ctx, int code, const std::string& msg) { return ctx_set_error(ctx, code, msg); }

#define DENS_HIPCHK(ctx, expr)                                                                                \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return fail((ctx), PSF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t kMaxCellsPerAxis = 1024;

DensityGrid grid_of(const psf_scene* s)
{
    DensityGrid D{};
    for (int a = 0; a < 3; ++a) {
        D.amin[a] = s->d.aabb_min[a];
        D.ext[a] = s->d.aabb_max[a] - s->d.aabb_min[a];   // AABB::getExtent, float
        D.n[a] = s->dens_n[a];
        D.r[a] = static_cast<float>(s->dens_n[a]);
    }
    D.counts = s->dev.dcounts;
    return D;
}

size_t cells_of(const psf_scene* s) { return size_t(s->dens_n[0]) * s->dens_n[1] * s->dens_n[2]; }

int zero_counts(psf_scene* s)
{
    DENS_HIPCHK(s->ctx, launch_zero_int(reinterpret_cast<int*>(s->dev.dcounts), static_cast<int>(s->dens_cells),
                                        s->ctx->stream));
    s->dens_state = 0;
    return PSF_OK;
}

}  // namespace

int density_for_trace(psf_scene* s, DensityGrid& D)
{
    D = DensityGrid{};
    if (!s->dev.dcounts) return PSF_OK;
    if (!s->dens_in_trace) {
        density_particles_changed(s);
        return PSF_OK;
    }
    if (s->dens_state != 0)
        if (int rc = zero_counts(s)) return rc;
    D = grid_of(s);
    s->dens_state = 1;
    return PSF_OK;
}

void density_particles_changed(psf_scene* s)
{
    if (s->dens_state == 1) s->dens_state = 2;
}

void density_release(psf_scene* s)
{
    scene_free(s, s->dev.dcounts);
    scene_free(s, s->dev.dvol);
    s->dev.dcounts = nullptr;
    s->dev.dvol = nullptr;
    s->dens_n[0] = s->dens_n[1] = s->dens_n[2] = 0;
    s->dens_cells = 0;
    s->dens_in_trace = false;
    s->dens_state = 0;
}

int density_enqueue_count(psf_scene* s)
{
    DENS_HIPCHK(s->ctx, launch_density_count(grid_of(s), s->dev.parts, s->n_particles, s->ctx->stream));
    s->dens_state = 2;
    return PSF_OK;
}

int density_enqueue_volume(psf_scene* s)
{
    DENS_HIPCHK(s->ctx, launch_density_volume(s->dev.dcounts, s->dev.dvol, s->dens_cells / 2048, s->ctx->stream));
    s->dens_state = 0;
    return PSF_OK;
}

// the density of the stored particles into the volume: a count pass unless
// the last particle pass counted them, then the volume pass
int density_enqueue(psf_scene* s, uint32_t* counts_host)
{
    if (int rc = scene_flush_traces(s)) return rc;   // the particles of every frame advanced so far
    if (s->dens_state != 1) {
        if (s->dens_state != 0)
            if (int rc = zero_counts(s)) return rc;
        if (int rc = density_enqueue_count(s)) return rc;
        s->dens_state = 1;
    }
    if (counts_host)
        DENS_HIPCHK(s->ctx, hipMemcpyAsync(counts_host, s->dev.dcounts, cells_of(s) * 4, hipMemcpyDeviceToHost,
                                           s->ctx->stream));
    return density_enqueue_volume(s);
}

}  // namespace psf

using namespace psf;

extern "C" {

int psf_set_density_grid(psf_scene* s, uint32_t rx, uint32_t ry, uint32_t rz, int count_in_trace)
{
    if (!s) return PSF_ERR_ARG;
    psf_ctx* ctx = s->ctx;
    const bool off = rx == 0 && ry == 0 && rz == 0;
    if (!off && (rx == 0 || ry == 0 || rz == 0 || rx > kMaxCellsPerAxis || ry > kMaxCellsPerAxis ||
                 rz > kMaxCellsPerAxis))
        return fail(ctx, PSF_ERR_ARG, "density grid: 1..1024 cells per axis, or 0,0,0 to remove it");
    DENS_HIPCHK(ctx, hipSetDevice(ctx->device));
    if (int rc = scene_flush_traces(s)) return rc;         // deferred passes count into the old grid
    DENS_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));   // earlier passes may still count into the old grid
    density_release(s);
    if (off) return PSF_OK;
    const size_t cells = (size_t(rx) * ry * rz + 2047) / 2048 * 2048;   // whole volume-pass blocks
    void* c = nullptr;
    void* v = nullptr;
    hipError_t e = scene_alloc(s, &c, cells * 4);
    if (e == hipSuccess) e = scene_alloc(s, &v, cells * 2);
    if (e != hipSuccess) {
        scene_free(s, c);
        return fail(ctx, e == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP,
                    std::string("density grid: ") + hipGetErrorString(e));
    }
    s->dev.dcounts = static_cast<uint32_t*>(c);
    s->dev.dvol = static_cast<uint16_t*>(v);
    s->dens_n[0] = rx;
    s->dens_n[1] = ry;
    s->dens_n[2] = rz;
    s->dens_cells = cells;
    s->dens_in_trace = count_in_trace != 0;
    if (int rc = zero_counts(s)) return rc;
    DENS_HIPCHK(ctx, hipMemsetAsync(s->dev.dvol, 0, cells * 2, ctx->stream));
    return PSF_OK;
}

int psf_density(psf_scene* s, uint32_t* counts, uint16_t* volume)
{
    if (!s) return PSF_ERR_ARG;
    psf_ctx* ctx = s->ctx;
    if (!s->dev.dcounts) return fail(ctx, PSF_ERR_STATE, "no density grid (psf_set_density_grid)");
    DENS_HIPCHK(ctx, hipSetDevice(ctx->device));
    if (int rc = scene_flush_traces(s)) return rc;
    if (int rc = density_enqueue(s, counts)) return rc;
    if (volume)
        DENS_HIPCHK(ctx, hipMemcpyAsync(volume, s->dev.dvol, cells_of(s) * 2, hipMemcpyDeviceToHost, ctx->stream));
    if (counts || volume) DENS_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PSF_OK;
}

}  // extern "C"
