// C ABI of the psf HIP library (declared in include/psf/psf.h).
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include <map>
#include <mutex>
#include <set>
#include <unordered_map>

#include "psf_internal.hpp"
#include "psf_kernels.hpp"
#include "psf_persistent.hpp"

using namespace psf;

namespace psf {

// Size classes (powers of two from 4 KiB) of device memory, recycled
// between the pooled scenes of one context; returned to HIP with the context.
struct DevicePool {
    std::multimap<size_t, void*> idle;
    std::unordered_map<void*, size_t> live;
    size_t held = 0;
};

static size_t pool_class(size_t bytes)
{
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

hipError_t scene_alloc(psf_scene* s, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    if (!s->pooled) return hipMalloc(p, bytes);
    DevicePool*& pool = s->ctx->pool;
    if (!pool) {
        pool = new (std::nothrow) DevicePool();
        if (!pool) return hipErrorOutOfMemory;
    }
    const size_t c = pool_class(bytes);
    auto it = pool->idle.find(c);
    if (it != pool->idle.end()) {
        *p = it->second;
        pool->idle.erase(it);
    } else {
        hipError_t e = hipMalloc(p, c);
        if (e != hipSuccess) return e;
        pool->held += c;
    }
    pool->live[*p] = c;
    return hipSuccess;
}

void scene_free(psf_scene* s, void* p)
{
    if (!p) return;
    if (!s->pooled || !s->ctx->pool) {
        (void)hipFree(p);
        return;
    }
    DevicePool* pool = s->ctx->pool;
    auto it = pool->live.find(p);
    if (it == pool->live.end()) {
        (void)hipFree(p);
        return;
    }
    pool->idle.emplace(it->second, p);
    pool->live.erase(it);
}

size_t pool_class_bytes(size_t bytes) { return bytes ? pool_class(bytes) : 0; }

static void pool_release(psf_ctx* c)
{
    if (!c->pool) return;
    for (auto& kv : c->pool->idle) (void)hipFree(kv.second);
    for (auto& kv : c->pool->live) (void)hipFree(kv.first);
    delete c->pool;
    c->pool = nullptr;
}

}  // namespace psf

// host storage of a scene's last viscosity factors (defined with psf_scene_host)
std::vector<double>& visc_of(psf_scene* s);

namespace {

// frames whose particle passes are deferred and launched together
constexpr uint32_t kTraceFuse = 8;

int set_err(psf_ctx* ctx, int code, const char* fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_err((ctx), PSF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

template <typename T>
hipError_t salloc(psf_scene* s, T** p, size_t n)
{
    return scene_alloc(s, reinterpret_cast<void**>(p), n * sizeof(T));
}

void sfree(psf_scene* s, void* p) { scene_free(s, p); }

void release_side(psf_scene* s)
{
    if (s->side) (void)hipStreamDestroy(s->side);
    for (int q = 0; q < 2; ++q)
        if (s->ev_ready[q]) (void)hipEventDestroy(s->ev_ready[q]);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    s->side = nullptr;
}

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

bool is_plain_precision(int p) { return p == PSF_C_FP64 || p == PSF_C_FP32; }

size_t c_bytes(const psf_scene_desc& d)
{
    const size_t K = d.K;
    switch (d.c_precision) {
    case PSF_C_FP64: return K * K * K * 8;
    case PSF_C_FP32: return K * K * K * 4;
    default: return persistent_c_bytes(d.c_precision, static_cast<int>(d.K));
    }
}

void counts(const psf_scene_desc& d, uint32_t& V, uint32_t& F, uint32_t& nZF, uint32_t& nYF, uint32_t& nXF)
{
    nXF = (d.nx + 1) * (d.ny * d.nz);
    nYF = (d.ny + 1) * (d.nx * d.nz);
    nZF = (d.nz + 1) * (d.nx * d.ny);
    F = nXF + nYF + nZF;
    V = d.nx * d.ny * d.nz;
}

int validate_desc(psf_ctx* ctx, const psf_scene_desc* d)
{
    if (!d) return set_err(ctx, PSF_ERR_ARG, "desc is null");
    if (d->nx == 0 || d->ny == 0 || d->nz == 0 || d->nx > 2048 || d->ny > 2048 || d->nz > 2048)
        return set_err(ctx, PSF_ERR_ARG, "grid dims out of range");
    const double nf = (double)(d->nx + 1) * d->ny * d->nz + (double)(d->ny + 1) * d->nx * d->nz +
                      (double)(d->nz + 1) * d->nx * d->ny;
    // the particle pass addresses the face field by 32-bit byte offsets
    if (nf >= 1073741824.0) return set_err(ctx, PSF_ERR_ARG, "too many faces (at most 2^30 - 1)");
    if (d->K == 0 || d->K > 4096) return set_err(ctx, PSF_ERR_ARG, "K must be in [1, 4096]");
    if (!(d->resolution > 0.0) || !std::isfinite(d->resolution)) return set_err(ctx, PSF_ERR_ARG, "resolution must be > 0");
    if (d->c_precision < PSF_C_FP64 || d->c_precision > PSF_C_BF16X2)
        return set_err(ctx, PSF_ERR_ARG, "unknown c_precision %d", d->c_precision);
    if (d->phi_precision != PSF_PHI_FP64 && d->phi_precision != PSF_PHI_FP32 && d->phi_precision != PSF_PHI_FP32_ACC32)
        return set_err(ctx, PSF_ERR_ARG, "unknown phi_precision %d", d->phi_precision);
    for (int c = 0; c < 3; ++c)
        if (!(d->aabb_max[c] > d->aabb_min[c])) return set_err(ctx, PSF_ERR_ARG, "aabb_max must exceed aabb_min");
    return PSF_OK;
}

int check_status(psf_scene* s)
{
    int st = 0;
    HIPCHK(s->ctx, hipMemcpyAsync(&st, s->dev.status, sizeof(int), hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    if (st == 1) return set_err(s->ctx, PSF_ERR_NONFINITE, "energy renormalisation saw a zero or non-finite energy");
    if (st) return set_err(s->ctx, PSF_ERR_HIP, "persistent step: the exchange of y between workgroups timed out");
    return PSF_OK;
}

TraceParams trace_params(const psf_scene* s)
{
    TraceParams p{};
    p.nx = s->d.nx; p.ny = s->d.ny; p.nz = s->d.nz;
    p.nZF = s->nZF; p.nYF = s->nYF;
    for (int c = 0; c < 3; ++c) {
        p.aabb_min[c] = s->d.aabb_min[c];
        p.aabb_max[c] = s->d.aabb_max[c];
        p.ext[c] = s->d.aabb_max[c] - s->d.aabb_min[c];
    }
    p.res = s->d.resolution;
    p.inv_res = 1.0 / s->d.resolution;
    p.hf = static_cast<float>(s->d.resolution);
    p.inv_h = static_cast<float>(1.0 / s->d.resolution);
    p.dtf = static_cast<float>(s->dt);
    p.lifetime = s->lifetime;
    p.center_z = s->d.aabb_min[2] + 0.5f * (s->d.aabb_max[2] - s->d.aabb_min[2]);
    p.seed = s->seed;
    p.frame = s->frame;
    int e = 0;
    p.pow2 = std::frexp(s->d.resolution, &e) == 0.5 && static_cast<double>(p.inv_h) == 1.0 / s->d.resolution;
    return p;
}

int ready_for_step(psf_scene* s)
{
    if (!s->has_C) return set_err(s->ctx, PSF_ERR_STATE, "advection tensor not uploaded");
    if (!s->has_omega) return set_err(s->ctx, PSF_ERR_STATE, "omega not set");
    if (!s->has_basis) return set_err(s->ctx, PSF_ERR_STATE, "basis not uploaded");
    return PSF_OK;
}

// Events of the launches of a timed region (psf_time_phase sets
// time_kernels): each launch records its own start and end (hipExtLaunchKernel,
// no marker packets between the kernels); null when not timing.
hipEvent_t pool_event(psf_scene* s)
{
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        s->ev_pool.push_back(e);
    }
    return s->ev_pool[s->ev_used++];
}

const psf::LaunchEvents* timed_events(psf_scene* s, psf::LaunchEvents& ev)
{
    if (!s->time_kernels) return nullptr;
    ev.start = pool_event(s);
    ev.stop = pool_event(s);
    return ev.start && ev.stop ? &ev : nullptr;
}

void timed_done(psf_scene* s, int kind, uint32_t units, const psf::LaunchEvents* ev)
{
    if (ev) s->timed.push_back(psf_scene::TimedLaunch{kind, units, ev->start, ev->stop});
}

int enqueue_steps(psf_scene* s, int n, double* hist = nullptr, int hist_ld = 0, hipStream_t st = nullptr)
{
    const int K = static_cast<int>(s->d.K);
    if (!st) st = s->ctx->stream;
    if (s->path == PSF_PATH_PERSISTENT) {
        psf::LaunchEvents ev;
        const psf::LaunchEvents* pe = n > 0 ? timed_events(s, ev) : nullptr;
        hipError_t e = launch_persistent_steps(s, n, st, hist, hist_ld, pe);
        if (e != hipSuccess) return set_err(s->ctx, PSF_ERR_HIP, "persistent step launch: %s", hipGetErrorString(e));
        timed_done(s, 0, static_cast<uint32_t>(n), pe);
        if (n > 0) s->cnt.step_launches++;
        return PSF_OK;
    }
    if (!is_plain_precision(s->d.c_precision))
        return set_err(s->ctx, PSF_ERR_UNSUPPORTED, "c_precision %d runs on the persistent path only", s->d.c_precision);
    for (int it = 0; it < n; ++it) {
        HIPCHK(s->ctx, launch_contract_plain(s->d.c_precision, s->dev.C, s->dev.omega, s->dev.y, K, st));
        HIPCHK(s->ctx, launch_update(s->dev.omega, s->dev.y, s->dev.visc, s->dev.gravity, K, s->dt, s->gravity_active,
                                     s->dev.status, hist, hist_ld, it, st));
        s->cnt.step_launches++;
    }
    return PSF_OK;
}

// buffers of the frame window: 0 is dev.hist / dev.u_hist, 1 dev.hist_b / dev.u_hist_b
double* win_hist(psf_scene* s, uint32_t b) { return b ? s->dev.hist_b : s->dev.hist; }
float* win_u(psf_scene* s, uint32_t b) { return b ? s->dev.u_hist_b : s->dev.u_hist; }

// u of the current frame
float* current_u(psf_scene* s)
{
    return s->win_cursor > 0 ? win_u(s, s->win_buf) + size_t(s->win_cursor - 1) * s->dev.ldF : s->dev.u;
}

int enqueue_reconstruct(psf_scene* s)
{
    psf::LaunchEvents ev;
    const psf::LaunchEvents* pe = timed_events(s, ev);
    HIPCHK(s->ctx, launch_reconstruct(s->d.phi_precision, s->dev.Phi, s->dev.ldF, s->F, static_cast<int>(s->d.K),
                                      s->dev.omega, s->dev.u, s->ctx->stream, pe));
    timed_done(s, 1, 1, pe);
    s->cnt.recon_launches++;
    return PSF_OK;
}

// Particle passes of deferred frames, one launch (the per-frame bits; the
// density grid, if it counts in the particle pass, counts the last frame).
int flush_traces(psf_scene* s)
{
    if (s->trace_pending == 0) return PSF_OK;
    TraceParams p = trace_params(s);
    DensityGrid D;
    if (int rc = density_for_trace(s, D)) return rc;
    psf::LaunchEvents ev;
    const psf::LaunchEvents* pe = timed_events(s, ev);
    HIPCHK(s->ctx, launch_trace_frames(p, win_u(s, s->win_buf) + size_t(s->trace_c0) * s->dev.ldF, s->dev.ldF,
                                       s->trace_pending, s->dev.parts, s->dev.pid, s->n_particles, D,
                                       s->ctx->stream, pe));
    timed_done(s, 2, s->trace_pending, pe);
    s->cnt.particle_launches++;
    s->frame += s->trace_pending;
    s->trace_pending = 0;
    return PSF_OK;
}

int enqueue_trace(psf_scene* s, int32_t* dcells)
{
    if (int rc = flush_traces(s)) return rc;
    TraceParams p = trace_params(s);
    DensityGrid D;
    if (int rc = density_for_trace(s, D)) return rc;
    psf::LaunchEvents ev;
    const psf::LaunchEvents* pe = timed_events(s, ev);
    HIPCHK(s->ctx, launch_trace(p, current_u(s), s->dev.parts, s->dev.pid, s->n_particles, dcells, D, s->ctx->stream,
                                pe));
    timed_done(s, 2, 1, pe);
    s->cnt.particle_launches++;
    s->frame++;
    return PSF_OK;
}

// n <= W reduced steps from dev.omega into columns 0..n-1 of buffer b's
// history, then their n velocity fields into buffer b's u rows, on stream st
int enqueue_window_fill(psf_scene* s, uint32_t b, hipStream_t st, uint32_t n)
{
    const int W = static_cast<int>(s->window);
    int rc = enqueue_steps(s, static_cast<int>(n), win_hist(s, b), W, st);
    if (rc) return rc;
    psf::LaunchEvents ev;
    const psf::LaunchEvents* pe = timed_events(s, ev);
    HIPCHK(s->ctx, launch_reconstruct_window(s->d.phi_precision, W, static_cast<int>(n), s->dev.Phi, s->dev.ldF, s->F,
                                             static_cast<int>(s->d.K), win_hist(s, b), win_u(s, b), st, pe));
    timed_done(s, 1, n, pe);
    s->cnt.recon_launches++;
    s->cnt.window_fills++;
    s->cnt.window_frames += n;
    return PSF_OK;
}

// frames the next window fill computes: the window, or in a bounded run no
// more than the frames still to come (0: none)
uint32_t next_fill_size(const psf_scene* s)
{
    return s->bounded ? std::min(s->window, s->fill_budget) : s->window;
}

// Ends a frame window early: w and u of the current frame become dev.omega and
// dev.u again, and the frames computed ahead are dropped.  Every call that
// changes the scene's state or runs a single phase calls this first.
int window_flush(psf_scene* s)
{
    if (int rc = flush_traces(s)) return rc;
    hipStream_t m = s->ctx->stream;
    uint64_t dropped = s->win_filled - s->win_cursor;
    if (s->win_ahead) {   // the side stream's window also moved dev.omega: wait for it
        HIPCHK(s->ctx, hipStreamWaitEvent(m, s->ev_ready[s->win_buf ^ 1u], 0));
        s->win_ahead = false;
        dropped += s->win_ahead_n;
    }
    if (dropped) {
        s->cnt.windows_dropped++;
        s->cnt.frames_dropped += dropped;
    }
    if (s->win_filled == 0) return PSF_OK;
    const int t = static_cast<int>(s->win_cursor) - 1;
    s->win_filled = 0;
    s->win_cursor = 0;
    if (t < 0) return PSF_OK;
    HIPCHK(s->ctx, launch_gather_column(win_hist(s, s->win_buf), static_cast<int>(s->window), t,
                                        static_cast<int>(s->d.K), s->dev.omega, m));
    HIPCHK(s->ctx, hipMemcpyAsync(s->dev.u, win_u(s, s->win_buf) + size_t(t) * s->dev.ldF, size_t(s->F) * 4,
                                  hipMemcpyDeviceToDevice, m));
    return PSF_OK;
}

// One integrate() frame: reduced step, u = Phi w, particle pass
// (psfsolver.cpp:21-234).  With a frame window of W > 1 the reduced steps and
// the reconstructions of W frames are computed together (the reduced dynamics
// do not depend on the particles), and each frame then only traces with its
// own u.  With the second buffer, the next window is computed on the side
// stream while the frames of the current one trace on the context stream.
int enqueue_frame(psf_scene* s)
{
    int rc;
    s->cnt.frames++;
    if (s->window <= 1) {
        if ((rc = enqueue_steps(s, 1))) return rc;
        if ((rc = enqueue_reconstruct(s))) return rc;
        return enqueue_trace(s, nullptr);
    }
    hipStream_t m = s->ctx->stream;
    if (s->win_cursor == s->win_filled) {
        uint32_t next = s->win_buf;
        if (s->win_filled > 0 && s->win_pipeline) next = s->win_buf ^ 1u;
        s->win_filled = 0;
        s->win_cursor = 0;
        uint32_t n;
        if (s->win_ahead) {
            HIPCHK(s->ctx, hipStreamWaitEvent(m, s->ev_ready[next], 0));
            s->win_ahead = false;
            n = s->win_ahead_n;
        } else {
            n = next_fill_size(s);
            if (n == 0) n = s->window;   // a bounded run asked for more frames than it announced
            if ((rc = enqueue_window_fill(s, next, m, n))) return rc;
            if (s->bounded) s->fill_budget -= std::min(n, s->fill_budget);
        }
        s->win_buf = next;
        s->win_filled = n;
        const uint32_t n2 = s->win_pipeline ? next_fill_size(s) : 0u;
        if (n2 > 0) {
            // the window after this one, on the side stream, behind everything
            // queued so far on the context stream (omega, the buffer's last traces)
            HIPCHK(s->ctx, hipEventRecord(s->ev_fork, m));
            HIPCHK(s->ctx, hipStreamWaitEvent(s->side, s->ev_fork, 0));
            if ((rc = enqueue_window_fill(s, next ^ 1u, s->side, n2))) return rc;
            HIPCHK(s->ctx, hipEventRecord(s->ev_ready[next ^ 1u], s->side));
            if (s->bounded) s->fill_budget -= n2;
            s->win_ahead = true;
            s->win_ahead_n = n2;
        }
    }
    s->win_cursor++;
    // defer this frame's particle pass; kTraceFuse frames (or the rest of
    // the window, whose u rows the next refill overwrites) go as one launch
    if (s->trace_pending == 0) s->trace_c0 = s->win_cursor - 1;
    s->trace_pending++;
    if (s->trace_pending == kTraceFuse || s->win_cursor == s->win_filled) return flush_traces(s);
    return PSF_OK;
}

// Host-to-device copy into a scene buffer on the context stream, so that it
// lands after every kernel queued there that reads or writes the buffer (and,
// through window_flush, after the side stream's window); returns when the
// copy is complete, so the host source may be freed.  (A null-stream copy is
// not ordered against the context stream, which is non-blocking.)
int upload_sync(psf_scene* s, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return PSF_OK;
    HIPCHK(s->ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return PSF_OK;
}

int refresh_visc(psf_scene* s, const std::vector<double>& lam)
{
    // omega_k *= exp(viscosity * eigenValues(k) * timeStep)  (psfsolver.cpp:51-55);
    // the factors stay in the scene's host storage, the copy's source
    std::vector<double>& v = visc_of(s);
    v.resize(s->d.K);
    for (uint32_t k = 0; k < s->d.K; ++k) v[k] = std::exp(s->nu * lam[k] * s->dt);
    HIPCHK(s->ctx, hipMemcpyAsync(s->dev.visc, v.data(), sizeof(double) * s->d.K, hipMemcpyHostToDevice, s->ctx->stream));
    if (!s->async_uploads) HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return PSF_OK;
}

}  // namespace

// host-side copies kept beside a scene: lambda (viscosity refreshes) and the
// caller index of every stored particle
struct psf_scene_host {
    std::vector<double> lambda;
    std::vector<uint32_t> pid;
    std::vector<double> visc;    // source of the last viscosity upload
};
static psf_scene_host* host_of(psf_scene* s) { return reinterpret_cast<psf_scene_host*>(s->host); }
std::vector<double>& visc_of(psf_scene* s) { return host_of(s)->visc; }

namespace psf {
// omega (dev.omega) and the particles in the caller's order, read with
// synchronous copies on the null stream (the scene's work must be complete);
// reports the scene's status word like psf_get_omega
int scene_read_results(psf_scene* s, double* omega, float* xyzw)
{
    int st = 0;
    HIPCHK(s->ctx, hipMemcpy(&st, s->dev.status, sizeof(int), hipMemcpyDeviceToHost));
    if (st == 1) return set_err(s->ctx, PSF_ERR_NONFINITE, "energy renormalisation saw a zero or non-finite energy");
    if (st) return set_err(s->ctx, PSF_ERR_HIP, "persistent step: the exchange of y between workgroups timed out");
    if (omega) HIPCHK(s->ctx, hipMemcpy(omega, s->dev.omega, s->d.K * 8, hipMemcpyDeviceToHost));
    const uint32_t n = s->n_particles;
    if (xyzw && n) {
        std::vector<float> sorted(size_t(n) * 4);
        HIPCHK(s->ctx, hipMemcpy(sorted.data(), s->dev.parts, size_t(n) * 16, hipMemcpyDeviceToHost));
        const std::vector<uint32_t>& pid = host_of(s)->pid;
        for (uint32_t j = 0; j < n; ++j) std::memcpy(xyzw + 4 * size_t(pid[j]), &sorted[4 * size_t(j)], 16);
    }
    return PSF_OK;
}

TraceParams scene_trace_params(const psf_scene* s) { return trace_params(s); }
int scene_flush_traces(psf_scene* s) { return flush_traces(s); }
int scene_window_flush(psf_scene* s) { return window_flush(s); }
int scene_check_status(psf_scene* s) { return check_status(s); }
int ctx_set_error(psf_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}
}  // namespace psf



namespace {

// Contexts alive in the process.  The first psf_ctx_create registers, with
// std::atexit, a handler that destroys the contexts a caller leaves alive at
// exit.  It registers right after the HIP runtime has answered, i.e. after
// the runtime's own initialisation registered its teardown, and exit runs
// handlers and static destructors in reverse order of registration: so the
// contexts' streams, events and memory are released while the runtime still
// holds them.  (A Python caller's own atexit hook, which runs earlier still,
// usually leaves nothing for it.)  Both containers are never destroyed, so
// no static destructor of this library runs after the runtime's teardown.
std::mutex& live_mutex()
{
    static std::mutex* m = new std::mutex();
    return *m;
}
std::set<psf_ctx*>& live_contexts()
{
    static std::set<psf_ctx*>* s = new std::set<psf_ctx*>();
    return *s;
}
void release_live_contexts()
{
    std::vector<psf_ctx*> left;
    {
        std::lock_guard<std::mutex> g(live_mutex());
        left.assign(live_contexts().begin(), live_contexts().end());
    }
    for (psf_ctx* c : left) psf_ctx_destroy(c);
}

}  // namespace

extern "C" {

int psf_abi_version(void) { return PSF_ABI_VERSION; }

int psf_ctx_create(int device, psf_ctx** out)
{
    if (!out) return PSF_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PSF_ERR_HIP;
    if (device < 0 || device >= n) return PSF_ERR_ARG;
    psf_ctx* c = new (std::nothrow) psf_ctx();
    if (!c) return PSF_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PSF_ERR_HIP;
    }
    static std::once_flag registered;
    std::call_once(registered, [] { std::atexit(release_live_contexts); });
    {
        std::lock_guard<std::mutex> g(live_mutex());
        live_contexts().insert(c);
    }
    *out = c;
    return PSF_OK;
}

int psf_ctx_destroy(psf_ctx* ctx)
{
    if (!ctx) return PSF_ERR_ARG;
    {
        std::lock_guard<std::mutex> g(live_mutex());
        live_contexts().erase(ctx);
    }
    (void)hipSetDevice(ctx->device);
    while (!ctx->servers.empty()) psf_server_destroy(*ctx->servers.begin());
    while (!ctx->scenes.empty()) psf_scene_destroy(*ctx->scenes.begin());
    pool_release(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PSF_OK;
}

const char* psf_last_error(const psf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int psf_ctx_sync(psf_ctx* ctx)
{
    if (!ctx) return PSF_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    for (psf_scene* s : ctx->scenes)
        if (int rc = flush_traces(s)) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (psf_scene* s : ctx->scenes)   // windows computed ahead on the side streams
        if (s->side) HIPCHK(ctx, hipStreamSynchronize(s->side));
    return PSF_OK;
}

void* psf_ctx_stream(psf_ctx* ctx) { return ctx ? reinterpret_cast<void*>(ctx->stream) : nullptr; }

int psf_scene_footprint(const psf_scene_desc* d, uint64_t* bytes)
{
    if (!bytes) return PSF_ERR_ARG;
    int rc = validate_desc(nullptr, d);
    if (rc) return rc;
    uint32_t V, F, a, b, c;
    counts(*d, V, F, a, b, c);
    const size_t ldF = round_up(F, 256);
    const size_t phi = ldF * d->K * (d->phi_precision == PSF_PHI_FP64 ? 8 : 4);
    *bytes = c_bytes(*d) + phi + size_t(F) * 4 + size_t(d->max_particles) * (16 + 48) + size_t(d->K) * 8 * 6 + 256;
    return PSF_OK;
}

int psf_scene_create(psf_ctx* ctx, const psf_scene_desc* desc, psf_scene** out)
{
    return scene_create_ex(ctx, desc, SceneOptions{}, out);
}

}  // extern "C"

int psf::scene_create_ex(psf_ctx* ctx, const psf_scene_desc* desc, const SceneOptions& opt, psf_scene** out)
{
    if (!ctx || !out) return set_err(ctx, PSF_ERR_ARG, "null argument");
    *out = nullptr;
    int rc = validate_desc(ctx, desc);
    if (rc) return rc;
    if (!is_plain_precision(desc->c_precision) &&
        !persistent_supported(desc->c_precision, static_cast<int>(desc->K), opt.modes_per_wg))
        return set_err(ctx, PSF_ERR_UNSUPPORTED, "c_precision %d with K=%u is not built", desc->c_precision, desc->K);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    psf_scene* s = new (std::nothrow) psf_scene();
    psf_scene_host* h = new (std::nothrow) psf_scene_host();
    if (!s || !h) {
        delete s;
        delete h;
        return set_err(ctx, PSF_ERR_NOMEM, "host allocation failed");
    }
    s->ctx = ctx;
    s->d = *desc;
    s->host = h;
    s->pooled = opt.pooled;
    s->modes_per_wg = opt.modes_per_wg;
    s->win_pipeline = opt.frame_pipeline;
    s->async_uploads = opt.async_uploads;
    counts(*desc, s->V, s->F, s->nZF, s->nYF, s->nXF);
    s->path = is_plain_precision(desc->c_precision) ? PSF_PATH_PLAIN : PSF_PATH_PERSISTENT;
    const size_t K = desc->K;
    s->dev.ldF = round_up(s->F, 256);
    s->dev.C_bytes = opt.no_c ? 0
                     : is_plain_precision(desc->c_precision)
                         ? c_bytes(*desc)
                         : persistent_c_bytes(desc->c_precision, static_cast<int>(K), opt.modes_per_wg);
    const size_t phi_bytes = s->dev.ldF * K * (desc->phi_precision == PSF_PHI_FP64 ? 8 : 4);
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    chk(salloc(s, &s->dev.omega, K));
    chk(salloc(s, &s->dev.lambda, K));
    chk(salloc(s, &s->dev.gravity, K));
    chk(salloc(s, &s->dev.visc, K));
    chk(salloc(s, &s->dev.y, K * static_cast<size_t>(plain_splits(static_cast<int>(K)))));   // plain-path partials
    chk(salloc(s, &s->dev.status, 64));
    chk(salloc(s, reinterpret_cast<unsigned char**>(&s->dev.C), s->dev.C_bytes));
    if (opt.borrowed_phi) {
        s->dev.Phi = opt.borrowed_phi;
        s->phi_borrowed = true;
        s->has_phi = true;
    } else {
        chk(salloc(s, reinterpret_cast<unsigned char**>(&s->dev.Phi), phi_bytes));
    }
    chk(salloc(s, &s->dev.u, s->F));
    chk(salloc(s, &s->dev.parts, desc->max_particles));
    if (opt.cells) chk(salloc(s, &s->dev.cells, size_t(desc->max_particles) * 12));
    chk(salloc(s, &s->dev.pid, desc->max_particles));
    if (e == hipSuccess && !opt.no_c) e = persistent_scene_init(s);
    if (e != hipSuccess) {
        psf_scene_destroy(s);
        return set_err(ctx, e == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP, "scene allocation: %s",
                       hipGetErrorString(e));
    }
    // zero padding of Phi (rows F..ldF-1 of each column are read by the
    // vectorised reconstruction) and the status word
    if (!s->phi_borrowed) chk(hipMemsetAsync(s->dev.Phi, 0, phi_bytes, ctx->stream));
    if (!opt.no_init) {
        chk(launch_zero_int(s->dev.status, 64, ctx->stream));
        // constants by fill kernels: host-to-device copies from pageable memory
        // measured 0.5 ms per scene in a long-running process (scene admission)
        chk(launch_fill_f64(s->dev.visc, static_cast<int>(K), 1.0, ctx->stream));
        chk(launch_fill_f64(s->dev.gravity, static_cast<int>(K), 0.0, ctx->stream));
        chk(launch_fill_f64(s->dev.lambda, static_cast<int>(K), 0.0, ctx->stream));
    }
    if (!s->async_uploads) chk(hipStreamSynchronize(ctx->stream));
    if (e != hipSuccess) {
        psf_scene_destroy(s);
        return set_err(ctx, PSF_ERR_HIP, "scene init: %s", hipGetErrorString(e));
    }
    h->lambda.assign(K, 0.0);
    ctx->scenes.insert(s);
    *out = s;
    return PSF_OK;
}

extern "C" {

int psf_scene_destroy(psf_scene* s)
{
    if (!s) return PSF_ERR_ARG;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->side) (void)hipStreamSynchronize(s->side);
    scene_release(s);
    return PSF_OK;
}

}  // extern "C"

// Frees a scene whose device work is known to be complete, without waiting
// on the context stream.
void psf::scene_release(psf_scene* s)
{
    persistent_scene_free(s);
    sfree(s, s->dev.omega); sfree(s, s->dev.lambda); sfree(s, s->dev.gravity); sfree(s, s->dev.visc);
    sfree(s, s->dev.y); sfree(s, s->dev.status); sfree(s, s->dev.C);
    if (!s->phi_borrowed) sfree(s, s->dev.Phi);
    sfree(s, s->dev.u); sfree(s, s->dev.parts); sfree(s, s->dev.cells); sfree(s, s->dev.pid);
    sfree(s, s->dev.hist); sfree(s, s->dev.u_hist); sfree(s, s->dev.omega_view);
    sfree(s, s->dev.hist_b); sfree(s, s->dev.u_hist_b);
    density_release(s);
    release_side(s);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    s->ctx->scenes.erase(s);
    delete host_of(s);
    delete s;
}

extern "C" {

int psf_scene_upload_basis(psf_scene* s, const double* Phi, const double* lambda, const double* gravity)
{
    if (!s || !lambda) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (int rc0 = window_flush(s)) return rc0;
    const size_t K = s->d.K, F = s->F, ld = s->dev.ldF;
    for (size_t k = 0; k < K; ++k)
        if (!std::isfinite(lambda[k])) return set_err(ctx, PSF_ERR_ARG, "lambda[%zu] is not finite", k);
    int rc = PSF_OK;
    if (Phi) {
        // column by column through a pinned staging buffer, converted to the storage type
        const size_t esz = s->d.phi_precision == PSF_PHI_FP64 ? 8 : 4;
        void* stage = nullptr;
        HIPCHK(ctx, hipHostMalloc(&stage, F * esz, hipHostMallocDefault));
        for (size_t k = 0; k < K && rc == PSF_OK; ++k) {
            const double* col = Phi + k * F;
            if (esz == 4) {
                float* sg = static_cast<float*>(stage);
                for (size_t f = 0; f < F; ++f) sg[f] = static_cast<float>(col[f]);
            } else {
                std::memcpy(stage, col, F * 8);
            }
            hipError_t e = hipMemcpyAsync(static_cast<unsigned char*>(s->dev.Phi) + k * ld * esz, stage, F * esz,
                                          hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = set_err(ctx, PSF_ERR_HIP, "Phi upload: %s", hipGetErrorString(e));
        }
        (void)hipHostFree(stage);
        if (rc == PSF_OK) s->has_phi = true;
    }
    if (rc) return rc;
    host_of(s)->lambda.assign(lambda, lambda + K);
    HIPCHK(ctx, hipMemcpyAsync(s->dev.lambda, lambda, K * 8, hipMemcpyHostToDevice, ctx->stream));
    if (gravity) {
        HIPCHK(ctx, hipMemcpyAsync(s->dev.gravity, gravity, K * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    if (!s->async_uploads) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    s->has_basis = true;
    return refresh_visc(s, host_of(s)->lambda);
}

int psf_scene_upload_advection(psf_scene* s, const double* C)
{
    if (!s || !C) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (int rc0 = window_flush(s)) return rc0;
    const size_t K = s->d.K, n = K * K * K;
    if (s->d.c_precision == PSF_C_FP64) {
        if (int rc = upload_sync(s, s->dev.C, C, n * 8)) return rc;
    } else if (s->d.c_precision == PSF_C_FP32) {
        std::vector<float> t(n);
        for (size_t i = 0; i < n; ++i) t[i] = static_cast<float>(C[i]);
        if (int rc = upload_sync(s, s->dev.C, t.data(), n * 4)) return rc;
    } else {
        int rc = persistent_upload_c(s, C);
        if (rc) return rc;
    }
    s->has_C = true;
    return PSF_OK;
}

int psf_scene_upload_advection_matrices(psf_scene* s, const double* const* mats)
{
    if (!s || !mats) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    const size_t K = s->d.K;
    std::vector<double> C;
    try {
        C.resize(K * K * K);
    } catch (const std::bad_alloc&) {
        return set_err(s->ctx, PSF_ERR_NOMEM, "host allocation of %zu doubles failed", K * K * K);
    }
    for (size_t i = 0; i < K; ++i) {
        if (!mats[i]) return set_err(s->ctx, PSF_ERR_ARG, "matrix %zu is null", i);
        std::memcpy(C.data() + i * K * K, mats[i], K * K * sizeof(double));
    }
    return psf_scene_upload_advection(s, C.data());
}

int psf_scene_check_signs(psf_scene* s, const uint32_t* words, size_t nwords)
{
    if (!s || !words) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    const size_t need = (s->V + 4) / 5;
    if (nwords != need) return set_err(s->ctx, PSF_ERR_ARG, "sign table has %zu words, expected %zu", nwords, need);
    std::vector<uint32_t> ref(need, 0u);
    uint32_t v = 0;
    for (uint32_t z = 0; z < s->d.nz; ++z)
        for (uint32_t y = 0; y < s->d.ny; ++y)
            for (uint32_t x = 0; x < s->d.nx; ++x, ++v) {
                const uint32_t bits = (z == 0 ? 1u : 0u) | 2u | (y == 0 ? 4u : 0u) | 8u | (x == 0 ? 16u : 0u) | 32u;
                ref[v / 5] |= bits << (6 * (v % 5));
            }
    for (size_t i = 0; i < need; ++i)
        if (ref[i] != words[i])
            return set_err(s->ctx, PSF_ERR_ARG, "sign word %zu is 0x%08x, the grid's closed form gives 0x%08x", i,
                           words[i], ref[i]);
    return PSF_OK;
}

int psf_set_params(psf_scene* s, double dt, double viscosity, int gravity_active, float lifetime)
{
    if (!s) return PSF_ERR_ARG;
    if (!std::isfinite(dt) || !std::isfinite(viscosity)) return set_err(s->ctx, PSF_ERR_ARG, "non-finite parameter");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    // a caller that passes the same values every frame (the reference's
    // integrate() reads them every frame) keeps its frames computed ahead
    if (dt == s->dt && viscosity == s->nu && (gravity_active ? 1 : 0) == s->gravity_active &&
        std::memcmp(&lifetime, &s->lifetime, sizeof(float)) == 0)
        return PSF_OK;
    if (int rc0 = window_flush(s)) return rc0;
    s->dt = dt;
    s->nu = viscosity;
    s->gravity_active = gravity_active ? 1 : 0;
    s->lifetime = lifetime;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    return refresh_visc(s, host_of(s)->lambda);
}

int psf_set_step_path(psf_scene* s, int path)
{
    if (!s) return PSF_ERR_ARG;
    if (path != PSF_PATH_PLAIN && path != PSF_PATH_PERSISTENT) return set_err(s->ctx, PSF_ERR_ARG, "unknown path %d", path);
    if (path == PSF_PATH_PLAIN && !is_plain_precision(s->d.c_precision))
        return set_err(s->ctx, PSF_ERR_UNSUPPORTED, "c_precision %d has no plain path", s->d.c_precision);
    if (path == PSF_PATH_PERSISTENT && !persistent_supported(s->d.c_precision, static_cast<int>(s->d.K)))
        return set_err(s->ctx, PSF_ERR_UNSUPPORTED, "c_precision %d has no persistent path", s->d.c_precision);
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (int rc0 = window_flush(s)) return rc0;
    s->path = path;
    return PSF_OK;
}

int psf_set_wait_limit(psf_scene* s, double seconds)
{
    if (!s) return PSF_ERR_ARG;
    if (!(seconds >= 0.0) || seconds > 3600.0) return set_err(s->ctx, PSF_ERR_ARG, "wait limit must be in [0, 3600] s");
    const uint64_t ticks = seconds == 0.0 ? 0ull : std::max<uint64_t>(1ull, static_cast<uint64_t>(seconds * 1e8));
    if (int rc = persistent_set_wait_ticks(s, ticks))
        return set_err(s->ctx, rc, "the scene has no persistent path (c_precision %d)", s->d.c_precision);
    return PSF_OK;
}

int psf_set_frame_window(psf_scene* s, uint32_t frames)
{
    if (!s) return PSF_ERR_ARG;
    if (frames != 1 && frames != 2 && frames != 4 && frames != 8 && frames != 16 && frames != 32)
        return set_err(s->ctx, PSF_ERR_ARG, "frame window must be 1, 2, 4, 8, 16 or 32 (got %u)", frames);
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (int rc0 = window_flush(s)) return rc0;
    if (frames == s->window) return PSF_OK;
    if (!s->async_uploads) {   // (a server's scene sets its window once, before any launch)
        HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
        if (s->side) HIPCHK(s->ctx, hipStreamSynchronize(s->side));
    }
    auto drop = [&]() {
        sfree(s, s->dev.hist); sfree(s, s->dev.u_hist); sfree(s, s->dev.omega_view);
        sfree(s, s->dev.hist_b); sfree(s, s->dev.u_hist_b);
        s->dev.hist = nullptr; s->dev.u_hist = nullptr; s->dev.omega_view = nullptr;
        s->dev.hist_b = nullptr; s->dev.u_hist_b = nullptr;
    };
    drop();
    s->window = 1;
    s->win_buf = 0;
    if (frames == 1) return PSF_OK;
    const size_t K = s->d.K;
    hipError_t e = salloc(s, &s->dev.hist, K * frames);
    if (e == hipSuccess) e = salloc(s, &s->dev.u_hist, s->dev.ldF * frames);
    if (e == hipSuccess) e = salloc(s, &s->dev.omega_view, K);
    if (s->win_pipeline) {
        if (e == hipSuccess) e = salloc(s, &s->dev.hist_b, K * frames);
        if (e == hipSuccess) e = salloc(s, &s->dev.u_hist_b, s->dev.ldF * frames);
        if (e == hipSuccess && !s->side) {
            e = hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking);
            for (int q = 0; q < 2 && e == hipSuccess; ++q)
                e = hipEventCreateWithFlags(&s->ev_ready[q], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming);
        }
    }
    if (e != hipSuccess) {
        drop();
        return set_err(s->ctx, e == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP, "frame window buffers: %s",
                       hipGetErrorString(e));
    }
    s->window = frames;
    return PSF_OK;
}

int psf_set_omega(psf_scene* s, const double* omega)
{
    if (!s || !omega) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (int rc0 = window_flush(s)) return rc0;
    HIPCHK(s->ctx, hipMemcpyAsync(s->dev.omega, omega, s->d.K * 8, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(s->ctx, launch_zero_int(s->dev.status, 64, s->ctx->stream));
    if (!s->async_uploads) HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    s->has_omega = true;
    return PSF_OK;
}

int psf_get_omega(psf_scene* s, double* omega)
{
    if (!s || !omega) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    const double* src = s->dev.omega;
    if (s->win_cursor > 0) {
        HIPCHK(s->ctx, launch_gather_column(win_hist(s, s->win_buf), static_cast<int>(s->window),
                                            static_cast<int>(s->win_cursor) - 1, static_cast<int>(s->d.K),
                                            s->dev.omega_view, s->ctx->stream));
        src = s->dev.omega_view;
    }
    HIPCHK(s->ctx, hipMemcpyAsync(omega, src, s->d.K * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    return check_status(s);
}

// Particles are stored sorted, stable in the caller's order.  A particle's
// interpolation reads twelve face pairs whose rows and planes are fixed by
// its half voxels on y and z (the M and O indices of both axes), and the
// texture path serves four consecutive lanes per cycle, one cache line each
// when they agree.  With at least 32 particles per half-voxel row on average
// (the y-z half voxels, 4 ny nz rows), they are sorted by (half voxel on z,
// half voxel on y, shifted cell on x): four consecutive particles then sit a
// few voxels apart in one row and read one line per gather.  Sparser sets
// would spread a row's few particles over the whole row, so they are sorted
// by the cell of the half-voxel-shifted grid (z-major), which keeps
// consecutive particles within a cell or two.  Every interface below speaks
// the caller's order.
}  // extern "C"

// The storage order of n particles in a scene of descriptor d: sorted[j] =
// xyzw[pid[j]], a stable counting sort by one of the two keys above.
void psf::sort_particles(const psf_scene_desc& d, const float* xyzw, uint32_t n, float* sorted, uint32_t* pid)
{
    std::iota(pid, pid + n, 0u);
    std::vector<uint32_t> key(n);
    const double res = d.resolution;
    const size_t rows = size_t(4) * d.ny * d.nz;
    const bool halves = size_t(n) >= 32 * rows;
    const size_t V = (halves ? rows : size_t(d.ny) * d.nz) * d.nx;
    const uint32_t dims[3] = {d.nx, d.ny, d.nz};
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            const float sa = (d.aabb_min[a] + (d.aabb_max[a] - d.aabb_min[a])) + xyzw[4 * size_t(i) + a];
            // the shifted cell (a / h - 1/2), or on y and z the half voxel (2 a / h)
            const bool half = halves && a > 0;
            const double q = half ? 2.0 * static_cast<double>(sa) / res : static_cast<double>(sa) / res - 0.5;
            const long hi = (half ? 2 : 1) * static_cast<long>(dims[a]) - 1;
            long v = std::isfinite(q) ? static_cast<long>(std::max(-1.0, std::min(q, 1.0e9))) : 0;
            c[a] = static_cast<uint32_t>(std::max(0L, std::min(v, hi)));
        }
        key[i] = (c[2] * (halves ? 2 * d.ny : d.ny) + c[1]) * d.nx + c[0];
    }
    if (V <= std::max<size_t>(size_t(4) * n, size_t(1) << 16)) {
        std::vector<uint32_t> start(V + 1, 0u);
        for (uint32_t i = 0; i < n; ++i) start[key[i] + 1]++;
        for (size_t v = 0; v < V; ++v) start[v + 1] += start[v];
        for (uint32_t i = 0; i < n; ++i) pid[start[key[i]]++] = i;
    } else {
        std::stable_sort(pid, pid + n, [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    }
    for (uint32_t j = 0; j < n; ++j) std::memcpy(sorted + 4 * size_t(j), xyzw + 4 * size_t(pid[j]), 16);
}

const std::vector<uint32_t>& psf::scene_pid(const psf_scene* s)
{
    return reinterpret_cast<const psf_scene_host*>(s->host)->pid;
}

void psf::scene_adopt_state(psf_scene* s, std::vector<uint32_t>&& pid, const double* lambda, double dt, double nu,
                            int gravity_active, float lifetime)
{
    psf_scene_host* h = host_of(s);
    s->n_particles = static_cast<uint32_t>(pid.size());
    h->pid = std::move(pid);
    h->lambda.assign(lambda, lambda + s->d.K);
    s->dt = dt;
    s->nu = nu;
    s->gravity_active = gravity_active ? 1 : 0;
    s->lifetime = lifetime;
    s->has_basis = true;
    s->has_omega = true;
    density_particles_changed(s);
}

extern "C" {

int psf_set_particles(psf_scene* s, const float* xyzw, uint32_t n)
{
    if (!s || (!xyzw && n)) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    if (n > s->d.max_particles) return set_err(s->ctx, PSF_ERR_ARG, "%u particles exceed max_particles %u", n, s->d.max_particles);
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (int rc = flush_traces(s)) return rc;   // deferred passes belong to the particles being replaced
    std::vector<uint32_t>& pid = host_of(s)->pid;
    pid.resize(n);
    std::vector<float> sorted(size_t(n) * 4);
    sort_particles(s->d, xyzw, n, sorted.data(), pid.data());
    if (n) {
        // after the passes just launched (and any earlier ones) on the context stream
        if (int rc = upload_sync(s, s->dev.parts, sorted.data(), size_t(n) * 16)) return rc;
        if (int rc = upload_sync(s, s->dev.pid, pid.data(), size_t(n) * 4)) return rc;
    }
    s->n_particles = n;
    density_particles_changed(s);
    return PSF_OK;
}

int psf_get_particles(psf_scene* s, float* xyzw, uint32_t n)
{
    if (!s || (!xyzw && n)) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    if (n != s->n_particles) return set_err(s->ctx, PSF_ERR_ARG, "asked for %u particles, scene holds %u", n, s->n_particles);
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (int rc = flush_traces(s)) return rc;
    std::vector<float> sorted(size_t(n) * 4);
    HIPCHK(s->ctx, hipMemcpyAsync(sorted.data(), s->dev.parts, size_t(n) * 16, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    const std::vector<uint32_t>& pid = host_of(s)->pid;
    for (uint32_t j = 0; j < n; ++j) std::memcpy(xyzw + 4 * size_t(pid[j]), &sorted[4 * size_t(j)], 16);
    return PSF_OK;
}

int psf_step(psf_scene* s, int nsteps)
{
    if (!s || nsteps < 0) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "bad argument");
    int rc = ready_for_step(s);
    if (rc) return rc;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if ((rc = window_flush(s))) return rc;
    return enqueue_steps(s, nsteps);
}

int psf_reconstruct(psf_scene* s)
{
    if (!s) return PSF_ERR_ARG;
    if (!s->has_phi || !s->has_omega) return set_err(s->ctx, PSF_ERR_STATE, "Phi or omega missing");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (int rc0 = window_flush(s)) return rc0;
    return enqueue_reconstruct(s);
}

int psf_get_velocity(psf_scene* s, float* u)
{
    if (!s || !u) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    HIPCHK(s->ctx, hipMemcpyAsync(u, current_u(s), size_t(s->F) * 4, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return PSF_OK;
}

int psf_trace(psf_scene* s, int32_t* cells)
{
    if (!s) return PSF_ERR_ARG;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    int rc = enqueue_trace(s, cells ? s->dev.cells : nullptr);
    if (rc || !cells) return rc;
    const uint32_t n = s->n_particles;
    std::vector<int32_t> sorted(size_t(n) * 12);
    HIPCHK(s->ctx, hipMemcpyAsync(sorted.data(), s->dev.cells, size_t(n) * 48, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    const std::vector<uint32_t>& pid = host_of(s)->pid;
    for (uint32_t j = 0; j < n; ++j) std::memcpy(cells + 12 * size_t(pid[j]), &sorted[12 * size_t(j)], 48);
    return PSF_OK;
}

int psf_integrate(psf_scene* s, int nframes)
{
    if (!s || nframes < 0) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "bad argument");
    int rc = ready_for_step(s);
    if (rc) return rc;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    if (!s->has_phi) return set_err(s->ctx, PSF_ERR_STATE, "Phi not uploaded");
    for (int f = 0; f < nframes; ++f)
        if ((rc = enqueue_frame(s))) return rc;
    return PSF_OK;
}

int psf_scene_device_view(psf_scene* s, psf_scene_view* out)
{
    if (!s || !out) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_scene_view v{};
    v.omega = s->dev.omega;
    v.lambda = s->dev.lambda;
    v.gravity = s->dev.gravity;
    v.visc = s->dev.visc;
    v.C = is_plain_precision(s->d.c_precision) ? s->dev.C : nullptr;
    v.Phi = s->dev.Phi;
    v.ld_phi = s->dev.ldF;
    v.u = current_u(s);
    v.particles = reinterpret_cast<float*>(s->dev.parts);
    v.pid = s->dev.pid;
    v.n_particles = s->n_particles;
    v.status = s->dev.status;
    *out = v;
    return PSF_OK;
}

int psf_scene_get_counters(psf_scene* s, psf_scene_counters* out)
{
    if (!s || !out) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    *out = s->cnt;
    return PSF_OK;
}

int psf_scene_region_kernels(psf_scene* s, psf_region_kernels* out)
{
    if (!s || !out) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    *out = s->region;
    return PSF_OK;
}

int psf_time_phase(psf_scene* s, int phase, int count, double* ms)
{
    if (!s || !ms || count <= 0 || phase < 0 || phase > 11) return set_err(s ? s->ctx : nullptr, PSF_ERR_ARG, "bad argument");
    // phase 11 is phase 3 with every launch of the region timed by its own
    // events (they lengthen the region by about a tenth, so phase 3 has none)
    const bool timed = phase == 11;
    if (timed) phase = 3;
    if (phase == 9 && (s->window < kTraceFuse || !s->has_phi))
        return set_err(s->ctx, PSF_ERR_STATE, "phase 9 needs Phi and a frame window of at least %u", kTraceFuse);
    if (phase == 5 && s->window <= 1) return set_err(s->ctx, PSF_ERR_STATE, "phase 5 needs a frame window > 1");
    if (phase >= 6 && phase <= 8 && !s->dev.dcounts)
        return set_err(s->ctx, PSF_ERR_STATE, "phases 6 to 8 need a density grid");
    if (phase != 2 && phase != 6 && phase != 7 && phase != 9) {
        int rc = ready_for_step(s);
        if (rc) return rc;
    }
    if ((phase == 1 || phase == 3 || phase == 5 || phase == 8 || phase == 10) && !s->has_phi)
        return set_err(s->ctx, PSF_ERR_STATE, "Phi not uploaded");
    psf_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // frames from a window boundary (3, 8) and the single phases start with
    // nothing computed ahead
    if (phase != 2 && phase != 6 && phase != 7 && phase != 9 && phase != 10) {
        int rc0 = window_flush(s);
        if (rc0) return rc0;
    }
    if (int rc0 = flush_traces(s)) return rc0;   // deferred passes belong before the timed region
    // everything queued so far, on either stream, ends before the start event
    if (phase == 3 || phase == 8) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (s->side) HIPCHK(ctx, hipStreamSynchronize(s->side));
    }
    hipEvent_t a, b;
    HIPCHK(ctx, hipEventCreate(&a));
    if (hipError_t eb = hipEventCreate(&b); eb != hipSuccess) {
        (void)hipEventDestroy(a);
        return set_err(ctx, PSF_ERR_HIP, "event create: %s", hipGetErrorString(eb));
    }
    // (no early return from here on: the flags below are cleared on every path)
    if (phase == 3 || phase == 8) {
        s->bounded = true;
        s->fill_budget = static_cast<uint32_t>(count);
    }
    if (timed) {
        // every kernel launched inside the region is timed on its stream
        s->time_kernels = true;
        s->timed.clear();
        s->ev_used = 0;
    }
    int rc = PSF_OK;
    hipError_t e = launch_timed_region_mark(0, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
    if (e != hipSuccess) rc = set_err(ctx, PSF_ERR_HIP, "event record: %s", hipGetErrorString(e));
    if (rc == PSF_OK) {
        switch (phase) {
        case 0: for (int i = 0; i < count && rc == PSF_OK; ++i) rc = enqueue_steps(s, 1); break;
        case 1: for (int i = 0; i < count && rc == PSF_OK; ++i) rc = enqueue_reconstruct(s); break;
        case 2: for (int i = 0; i < count && rc == PSF_OK; ++i) rc = enqueue_trace(s, nullptr); break;
        case 3:
        case 10: for (int i = 0; i < count && rc == PSF_OK; ++i) rc = enqueue_frame(s); break;
        case 5:
            for (int i = 0; i < count && rc == PSF_OK; ++i) {
                const int W = static_cast<int>(s->window);
                hipError_t le = launch_reconstruct_window(s->d.phi_precision, W, W, s->dev.Phi, s->dev.ldF, s->F,
                                                          static_cast<int>(s->d.K), s->dev.hist, s->dev.u_hist,
                                                          ctx->stream);   // buffer 0, after the flush
                if (le != hipSuccess) rc = set_err(ctx, PSF_ERR_HIP, "window reconstruction: %s", hipGetErrorString(le));
            }
            break;
        case 6: for (int i = 0; i < count && rc == PSF_OK; ++i) rc = density_enqueue_count(s); break;
        case 7: for (int i = 0; i < count && rc == PSF_OK; ++i) rc = density_enqueue_volume(s); break;
        case 8:
            for (int i = 0; i < count && rc == PSF_OK; ++i) {
                rc = enqueue_frame(s);
                if (rc == PSF_OK) rc = density_enqueue(s, nullptr);
            }
            break;
        case 9:
            // the frame loop's fused particle launch: kTraceFuse frames over
            // rows 0.. of the current window buffer
            for (int i = 0; i < count && rc == PSF_OK; ++i) {
                s->trace_c0 = 0;
                s->trace_pending = kTraceFuse;
                rc = flush_traces(s);
            }
            break;
        default: rc = enqueue_steps(s, count); break;
        }
    }
    s->bounded = false;
    s->fill_budget = 0;
    if (rc == PSF_OK) rc = flush_traces(s);      // the timed frames' deferred passes
    if (rc == PSF_OK && (phase == 3 || phase == 8) && s->win_ahead) {
        // (a bounded run computes nothing ahead; should it have, that work is in the region)
        e = hipStreamWaitEvent(ctx->stream, s->ev_ready[s->win_buf ^ 1u], 0);
        if (e != hipSuccess) rc = set_err(ctx, PSF_ERR_HIP, "stream wait: %s", hipGetErrorString(e));
    }
    if (rc == PSF_OK) {
        e = hipEventRecord(b, ctx->stream);
        if (e == hipSuccess) e = launch_timed_region_mark(1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
        if (e != hipSuccess) rc = set_err(ctx, PSF_ERR_HIP, "event timing: %s", hipGetErrorString(e));
        *ms = t;
    }
    if (rc == PSF_OK && s->time_kernels) {
        // the region's launches are complete: the end event follows them on
        // the context stream, which waited for the side stream's window
        psf_region_kernels r{};
        for (const psf_scene::TimedLaunch& l : s->timed) {
            float lt = 0.0f;
            e = hipEventSynchronize(l.b);
            if (e == hipSuccess) e = hipEventElapsedTime(&lt, l.a, l.b);
            if (e != hipSuccess) {
                rc = set_err(ctx, PSF_ERR_HIP, "kernel timing: %s", hipGetErrorString(e));
                break;
            }
            if (l.kind == 0) {
                r.step_ms += lt;
                r.step_launches++;
                r.steps += l.units;
            } else if (l.kind == 1) {
                r.recon_ms += lt;
                r.recon_launches++;
                r.recon_frames += l.units;
            } else {
                r.particle_ms += lt;
                r.particle_launches++;
                r.particle_frames += l.units;
            }
        }
        s->region = r;
    }
    s->time_kernels = false;
    s->timed.clear();
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc == PSF_OK && (phase == 0 || phase == 3 || phase == 4 || phase == 8 || phase == 10)) rc = check_status(s);
    return rc;
}

}  // extern "C"
