// Scene server: continuous batching of independent scenes (C ABI in psf.h).
//
// Every round advances all active scenes by up to `window` integrate() frames
// (psfsolver.cpp:21-234) with three launches on three streams:
//   1. P, on the context stream: one persistent launch holding every scene's
//      workgroups; each scene runs its own number of reduced steps and
//      writes its omega history;
//   2. R, on the reconstruction stream, after 1: one window reconstruction
//      over all scenes' Phi (FP64 sums) into their u history;
//   3. T, on the particle stream, after 2: one particle launch carrying
//      every scene's particles through their frames of the round.
// Each scene has two omega-history and two u-history buffers, used by rounds
// of alternate parity.  So round r+1's P and R run while round r's T does:
// P(r) waits for R(r-2) to release its omega history, R(r) for T(r-2) to
// release its u history, and T(r) follows T(r-1) on its stream.  P is
// latency-bound (one device-wide exchange per step), R FP64-bound and T
// gather-bound, so overlapping them fills the device better than running
// them in sequence.  The host keeps one round in flight: a call launches
// round r, then waits for round r-1 and hands out the requests that finished
// in it.
// Scenes never exchange data, and each scene's arithmetic depends only on its
// own inputs and layout, so a request gives the same bits alone and in a mix.
// Admission is FIFO at round start, bounded by active slots, the workgroups
// the persistent launch can hold resident, and a device-memory budget.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "psf_group.hpp"
#include "psf_internal.hpp"
#include "psf_kernels.hpp"
#include "psf_persistent.hpp"

using namespace psf;
using Clock = std::chrono::steady_clock;

namespace {

constexpr uint32_t kMaxActive = 64;
constexpr int kMaxGroupInstances = 16;   // shared-basis instances per launch

struct Basis {
    psf_scene_desc desc{};       // grid; K0 = desc.K
    uint32_t F = 0;
    size_t ldF = 0;
    float* Phi = nullptr;        // ldF x K0, FP32, rows F..ldF-1 zero
    double* C = nullptr;         // K0^3, FP64
    double* gravity_dev = nullptr;   // K0 (shared-basis steps)
    GroupBasis* group = nullptr;     // C's register blocks (shared-basis steps)
    std::vector<double> lambda, gravity;
};

enum ReqState { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3 };

struct Request {
    uint64_t id = 0;
    uint32_t basis = 0, K = 0, nframes = 0, sched = 0, n = 0;
    int64_t last_round = -1;     // round that runs its last frames
    double* hist2 = nullptr;     // second omega-history buffer (the first is the scene's)
    float* u_hist2 = nullptr;    // second u-history buffer
    double dt = 0.0, nu = 0.0;
    int grav = 0;
    float lifetime = 0.0f;
    std::vector<double> omega;   // input, then result
    std::vector<float> parts;    // result (caller's order)
    std::vector<float> sorted;   // input particles in storage order (sorted at submit)
    std::vector<uint32_t> pid;   // caller index of each stored particle
    void* pin = nullptr;         // pinned staging: the admission's uploads, then the results read back
    size_t pin_class = 0;
    psf_scene* scene = nullptr;
    uint64_t footprint = 0;
    int wgs = 0, ipw = 0;
    int state = kQueued;
    std::string error;
    Clock::time_point t_submit, t_done;
};

}  // namespace

// Recycled pinned host buffers in power-of-two size classes (admission
// staging): hipHostMalloc is slow, and host-to-device copies from pageable
// memory can block the host.
struct PinnedPool {
    std::multimap<size_t, void*> idle;
    size_t held = 0;
    ~PinnedPool()
    {
        for (auto& kv : idle) (void)hipHostFree(kv.second);
    }
    static size_t cls(size_t bytes)
    {
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        return c;
    }
    void* get(size_t bytes, size_t& c)
    {
        c = cls(bytes);
        auto it = idle.find(c);
        if (it != idle.end()) {
            void* p = it->second;
            idle.erase(it);
            return p;
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
        held += c;
        return p;
    }
    void put(void* p, size_t c)
    {
        if (p) idle.emplace(c, p);
    }
};

struct psf_server {
    psf_ctx* ctx = nullptr;
    PinnedPool pinned;
    psf_server_config cfg{};
    std::vector<Basis> bases;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::deque<uint64_t> finished;   // ids of requests done or failed, in that order (psf_server_completed)
    std::string last_diag;           // the last exchange wait that gave up (basis-shared steps)
    uint64_t next_id = 1;
    // per round parity: launch tables (pinned staging and device copies)
    PersistentBatch* batch[2] = {nullptr, nullptr};
    GroupBatch* gbatch[2] = {nullptr, nullptr};   // shared-basis steps
    WindowScene* d_wtab[2] = {nullptr, nullptr};
    WindowScene* h_wtab[2] = {nullptr, nullptr};
    TraceScene* d_ttab[2] = {nullptr, nullptr};
    TraceScene* h_ttab[2] = {nullptr, nullptr};
    hipEvent_t ev_step[2] = {nullptr, nullptr};   // after the persistent launch of a round
    hipEvent_t ev_rec[2] = {nullptr, nullptr};    // after the reconstruction of a round
    hipEvent_t ev_done[2] = {nullptr, nullptr};   // after the particle launch of a round
    hipStream_t second = nullptr;                 // particles
    hipStream_t third = nullptr;                  // reconstruction
    hipStream_t upload = nullptr;                 // admission uploads
    hipEvent_t ev_upload = nullptr;               // after the last admission's uploads
    bool upload_pending = false;
    int64_t launched = 0;                         // rounds launched
    int64_t retired = -1;                         // last round whose requests were handed out
    int wg_capacity = 0, wgs_in_use = 0;
    uint64_t mem_in_use = 0, mem_peak = 0;
    psf_server_stats st{};
};

namespace {

int fail(psf_ctx* ctx, int code, const std::string& msg) { return ctx_set_error(ctx, code, msg); }

#define SRV_HIPCHK(ctx, expr)                                                                                \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return fail((ctx), PSF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// bound of each exchange wait in s_memrealtime ticks (100 MHz; 0: the kernels' default)
uint64_t wait_ticks(const psf_server* sv)
{
    const uint64_t ns = sv->cfg.wait_limit_ns;
    return ns ? std::max<uint64_t>(1, ns / 10) : 0;
}

// modes per workgroup of a request's persistent layout
int request_ipw(const psf_server* sv, uint32_t K)
{
    return sv->cfg.modes_per_wg ? static_cast<int>(sv->cfg.modes_per_wg)
                                : persistent_auto_ipw(PSF_C_BF16X2, static_cast<int>(K));
}

// Device bytes of a request, as the pool holds them.
uint64_t request_bytes(const psf_server* sv, const Basis& b, uint32_t K, uint32_t n)
{
    // a shared-basis request holds no C of its own (the basis holds it once)
    const bool own_c = !sv->cfg.shared_basis;
    const int ipw = request_ipw(sv, K);
    const size_t cb = own_c ? persistent_c_bytes(PSF_C_BF16X2, static_cast<int>(K), ipw) : 0;
    const int G = own_c ? persistent_workgroups(PSF_C_BF16X2, static_cast<int>(K), ipw) : 0;
    const size_t bpw = G ? cb / (size_t(G) * 1024) : 0;
    const size_t Wn = sv->cfg.window;
    const size_t sizes[] = {
        K * 8ull, K * 8ull, K * 8ull, K * 8ull, K * 8ull, 64 * 4ull,   // omega lambda gravity visc y status
        cb, b.F * 4ull, n * 16ull, n * 4ull,                            // C u parts pid
        bpw * 4, G ? size_t(G) * ((K + G - 1) / G) * 8 : 0,              // meta inv_scale
        own_c ? 2ull * 16 * K * 8 : 0,                                  // xbuf
        K * Wn * 8, b.ldF * Wn * 4, K * 8ull, K * Wn * 8,               // hist u_hist omega_view hist2
        b.ldF * Wn * 4,                                                  // u_hist2
    };
    uint64_t t = 0;
    for (size_t x : sizes) t += pool_class_bytes(x);
    return t;
}

int admit(psf_server* sv, Request* r)
{
    const Basis& b = sv->bases[r->basis];
    psf_scene_desc d = b.desc;
    d.K = r->K;
    d.c_precision = PSF_C_BF16X2;
    d.phi_precision = sv->cfg.u_fp32_sums ? PSF_PHI_FP32_ACC32 : PSF_PHI_FP32;
    d.max_particles = r->n;
    SceneOptions o;
    o.modes_per_wg = r->ipw;
    o.borrowed_phi = b.Phi;
    o.cells = false;
    o.pooled = true;
    o.frame_pipeline = false;   // the server pipelines rounds itself
    o.async_uploads = true;     // P waits for the admission's copies on ev_upload
    o.no_c = sv->cfg.shared_basis != 0;   // steps read the basis's C
    o.no_init = true;           // every constant comes with the copies below
    psf_scene* s = nullptr;
    int rc = scene_create_ex(sv->ctx, &d, o, &s);
    if (rc) return rc;
    r->scene = s;
    if ((rc = psf_set_frame_window(s, sv->cfg.window))) return rc;
    SRV_HIPCHK(sv->ctx, scene_alloc(s, reinterpret_cast<void**>(&r->hist2), sizeof(double) * r->K * sv->cfg.window));
    SRV_HIPCHK(sv->ctx, scene_alloc(s, reinterpret_cast<void**>(&r->u_hist2), sizeof(float) * s->dev.ldF * sv->cfg.window));
    // Every upload from one pinned staging buffer, in copies that never wait
    // on the host: the particles in storage order and their caller indices,
    // omega, lambda, gravity, viscosity factors exp(nu lambda_k dt)
    // (psfsolver.cpp:51-55) and a cleared status word.
    const size_t K = r->K, n = r->n;
    const size_t o_pid = n * 16, o_w = o_pid + ((n * 4 + 15) & ~size_t(15)), o_l = o_w + K * 8, o_g = o_l + K * 8,
                 o_v = o_g + K * 8, o_st = o_v + K * 8, total = o_st + 256;
    r->pin = sv->pinned.get(total, r->pin_class);
    if (!r->pin) return fail(sv->ctx, PSF_ERR_NOMEM, "pinned staging of an admission");
    unsigned char* pin = static_cast<unsigned char*>(r->pin);
    std::memcpy(pin, r->sorted.data(), n * 16);
    std::memcpy(pin + o_pid, r->pid.data(), n * 4);
    std::memcpy(pin + o_w, r->omega.data(), K * 8);
    std::memcpy(pin + o_l, b.lambda.data(), K * 8);
    std::memcpy(pin + o_g, b.gravity.data(), K * 8);
    double* visc = reinterpret_cast<double*>(pin + o_v);
    for (size_t k = 0; k < K; ++k) visc[k] = std::exp(r->nu * b.lambda[k] * r->dt);
    std::memset(pin + o_st, 0, 256);
    hipStream_t up = sv->ctx->stream;   // the upload stream during admission
    auto h2d = [&](void* dst, size_t off, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, pin + off, bytes, hipMemcpyHostToDevice, up) : hipSuccess;
    };
    SRV_HIPCHK(sv->ctx, h2d(s->dev.parts, 0, n * 16));
    SRV_HIPCHK(sv->ctx, h2d(s->dev.pid, o_pid, n * 4));
    SRV_HIPCHK(sv->ctx, h2d(s->dev.omega, o_w, K * 8));
    SRV_HIPCHK(sv->ctx, h2d(s->dev.lambda, o_l, K * 8));
    SRV_HIPCHK(sv->ctx, h2d(s->dev.gravity, o_g, K * 8));
    SRV_HIPCHK(sv->ctx, h2d(s->dev.visc, o_v, K * 8));
    SRV_HIPCHK(sv->ctx, h2d(s->dev.status, o_st, 256));
    if (!sv->cfg.shared_basis) {
        SRV_HIPCHK(sv->ctx, launch_pack_c(s, b.C, static_cast<int>(b.desc.K), up));
        if (sv->cfg.wait_limit_ns) (void)persistent_set_wait_ticks(s, wait_ticks(sv));
    }
    s->has_C = true;
    scene_adopt_state(s, std::move(r->pid), b.lambda.data(), r->dt, r->nu, r->grav, r->lifetime);
    std::vector<float>().swap(r->sorted);
    return PSF_OK;
}

// omega, particles (caller's order) and status of a request whose last round
// has completed, from the copies its last round queued into its pinned buffer
int read_results(psf_server* sv, Request* r)
{
    const unsigned char* pin = static_cast<const unsigned char*>(r->pin);
    const size_t np = size_t(r->n) * 16, nw = size_t(r->K) * 8;
    int st = 0;
    std::memcpy(&st, pin + np + nw, sizeof(int));
    if (st == 1) return fail(sv->ctx, PSF_ERR_NONFINITE, "energy renormalisation saw a zero or non-finite energy");
    if (st) return fail(sv->ctx, PSF_ERR_HIP, "persistent step: the exchange of y between workgroups timed out");
    std::memcpy(r->omega.data(), pin + np, nw);
    const std::vector<uint32_t>& pid = scene_pid(r->scene);
    const float* sorted = reinterpret_cast<const float*>(pin);
    for (uint32_t j = 0; j < r->n; ++j) std::memcpy(r->parts.data() + 4 * size_t(pid[j]), sorted + 4 * size_t(j), 16);
    return PSF_OK;
}

// frees a request's scene; its device work must be complete
void release(psf_server* sv, Request* r)
{
    if (r->scene) {
        scene_free(r->scene, r->hist2);
        scene_free(r->scene, r->u_hist2);
        scene_release(r->scene);
    }
    r->scene = nullptr;
    r->hist2 = nullptr;
    r->u_hist2 = nullptr;
    sv->pinned.put(r->pin, r->pin_class);   // its copies completed with the request's rounds
    r->pin = nullptr;
    if (r->last_round < 0) sv->wgs_in_use -= r->wgs;   // otherwise freed when its last round launched
    sv->mem_in_use -= r->footprint;
}

// The scene calls used by admit() queue their copies and kernels on the
// context's stream; during admission that is the server's upload stream, so
// admission never waits for a running round.  The next persistent launch
// waits for the uploads through ev_upload.
struct UploadStream {
    psf_ctx* ctx;
    hipStream_t saved;
    UploadStream(psf_server* sv) : ctx(sv->ctx), saved(sv->ctx->stream) { ctx->stream = sv->upload; }
    ~UploadStream() { ctx->stream = saved; }
};

// Shared-basis steps: whether the next round's launch holds every active
// scene with frames left plus r (instances of up to kGroupMaxScenes scenes
// per basis, their workgroups resident at once).
bool group_fits(const psf_server* sv, const Request* r)
{
    std::vector<int> cnt(sv->bases.size(), 0);
    for (const Request* a : sv->active)
        if (a->sched < a->nframes) cnt[a->basis]++;
    cnt[r->basis]++;
    int wgs = 0, inst = 0;
    for (size_t b = 0; b < cnt.size(); ++b) {
        if (!cnt[b]) continue;
        const int ni = (cnt[b] + kGroupMaxScenes - 1) / kGroupMaxScenes;
        inst += ni;
        wgs += ni * group_basis_workgroups(sv->bases[b].group);
    }
    return wgs <= sv->wg_capacity && inst <= group_batch_max_instances(sv->gbatch[0]);
}

// Host time one round call may spend admitting while a round runs: longer
// admission would leave the device idle once the round in flight ends.
constexpr double kAdmitBudgetMs = 0.25;

void admit_queued(psf_server* sv, bool device_idle)
{
    UploadStream up(sv);
    const Clock::time_point t0 = Clock::now();
    while (!sv->queue.empty() && sv->active.size() < sv->cfg.max_active) {
        if (!device_idle && std::chrono::duration<double, std::milli>(Clock::now() - t0).count() > kAdmitBudgetMs)
            break;
        Request* r = sv->queue.front();
        if (sv->wgs_in_use + r->wgs > sv->wg_capacity) break;
        if (sv->cfg.shared_basis && !group_fits(sv, r)) break;
        if (sv->mem_in_use + r->footprint > sv->cfg.mem_budget) break;
        sv->queue.pop_front();
        sv->wgs_in_use += r->wgs;
        sv->mem_in_use += r->footprint;
        sv->mem_peak = std::max(sv->mem_peak, sv->mem_in_use);
        if (admit(sv, r) != PSF_OK) {
            r->error = sv->ctx->err;
            (void)hipStreamSynchronize(sv->upload);   // uploads already queued for the scene
            r->state = kFailed;
            r->t_done = Clock::now();
            sv->finished.push_back(r->id);
            release(sv, r);
            sv->st.failed++;
            continue;
        }
        r->state = kActive;
        sv->active.push_back(r);
        sv->upload_pending = true;
    }
    if (sv->upload_pending) (void)hipEventRecord(sv->ev_upload, sv->upload);
}

}  // namespace

extern "C" {

int psf_server_create(psf_ctx* ctx, const psf_server_config* cfg, psf_server** out)
{
    if (!ctx || !out) return PSF_ERR_ARG;
    *out = nullptr;
    psf_server_config c = cfg ? *cfg : psf_server_config{};
    if (c.window == 0) c.window = 32;
    if (c.max_active == 0) c.max_active = 32;
    if (c.mem_budget == 0) c.mem_budget = 64ull << 30;
    if (c.window != 2 && c.window != 4 && c.window != 8 && c.window != 16 && c.window != 32)
        return fail(ctx, PSF_ERR_ARG, "server window must be 2, 4, 8, 16 or 32");
    if (c.max_active > kMaxActive) return fail(ctx, PSF_ERR_ARG, "max_active must be at most 64");
    if (c.modes_per_wg > 8) return fail(ctx, PSF_ERR_ARG, "modes_per_wg must be at most 8");
    if (c.u_fp32_sums > 1) return fail(ctx, PSF_ERR_ARG, "u_fp32_sums must be 0 or 1");
    if (c.shared_basis > 1) return fail(ctx, PSF_ERR_ARG, "shared_basis must be 0 or 1");
    SRV_HIPCHK(ctx, hipSetDevice(ctx->device));
    psf_server* sv = new (std::nothrow) psf_server();
    if (!sv) return PSF_ERR_NOMEM;
    sv->ctx = ctx;
    sv->cfg = c;
    hipError_t e = hipStreamCreateWithFlags(&sv->second, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sv->third, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sv->upload, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->ev_upload, hipEventDisableTiming);
    for (int q = 0; q < 2 && e == hipSuccess; ++q) {
        if (c.shared_basis) {
            sv->gbatch[q] = group_batch_create(ctx->device, kMaxGroupInstances);
            if (!sv->gbatch[q]) e = hipErrorOutOfMemory;
        } else {
            sv->batch[q] = persistent_batch_create(ctx->device, static_cast<int>(kMaxActive), 8192);
            if (!sv->batch[q]) e = hipErrorOutOfMemory;
        }
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sv->d_wtab[q]), sizeof(WindowScene) * kMaxActive);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sv->d_ttab[q]), sizeof(TraceScene) * kMaxActive);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sv->h_wtab[q]), sizeof(WindowScene) * kMaxActive,
                              hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sv->h_ttab[q]), sizeof(TraceScene) * kMaxActive,
                              hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->ev_step[q], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->ev_rec[q], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->ev_done[q], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        ctx->servers.insert(sv);
        psf_server_destroy(sv);
        return fail(ctx, PSF_ERR_HIP, std::string("server set-up: ") + hipGetErrorString(e));
    }
    sv->wg_capacity = c.shared_basis ? group_batch_capacity(sv->gbatch[0])
                                     : persistent_batch_wg_capacity(sv->batch[0], PSF_C_BF16X2);
    ctx->servers.insert(sv);
    *out = sv;
    return PSF_OK;
}

int psf_server_destroy(psf_server* sv)
{
    if (!sv) return PSF_ERR_ARG;
    psf_ctx* ctx = sv->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (sv->second) (void)hipStreamSynchronize(sv->second);
    if (sv->third) (void)hipStreamSynchronize(sv->third);
    if (sv->upload) (void)hipStreamSynchronize(sv->upload);
    for (auto& kv : sv->reqs) {
        if (kv.second->scene) {
            scene_free(kv.second->scene, kv.second->hist2);
            scene_free(kv.second->scene, kv.second->u_hist2);
            scene_release(kv.second->scene);
        }
        sv->pinned.put(kv.second->pin, kv.second->pin_class);   // freed with the pool
    }
    for (Basis& b : sv->bases) {
        group_basis_destroy(b.group);
        if (b.Phi) (void)hipFree(b.Phi);
        if (b.C) (void)hipFree(b.C);
        if (b.gravity_dev) (void)hipFree(b.gravity_dev);
    }
    for (int q = 0; q < 2; ++q) {
        persistent_batch_destroy(sv->batch[q]);
        group_batch_destroy(sv->gbatch[q]);
        if (sv->d_wtab[q]) (void)hipFree(sv->d_wtab[q]);
        if (sv->d_ttab[q]) (void)hipFree(sv->d_ttab[q]);
        if (sv->h_wtab[q]) (void)hipHostFree(sv->h_wtab[q]);
        if (sv->h_ttab[q]) (void)hipHostFree(sv->h_ttab[q]);
        if (sv->ev_step[q]) (void)hipEventDestroy(sv->ev_step[q]);
        if (sv->ev_rec[q]) (void)hipEventDestroy(sv->ev_rec[q]);
        if (sv->ev_done[q]) (void)hipEventDestroy(sv->ev_done[q]);
    }
    if (sv->second) (void)hipStreamDestroy(sv->second);
    if (sv->third) (void)hipStreamDestroy(sv->third);
    if (sv->upload) (void)hipStreamDestroy(sv->upload);
    if (sv->ev_upload) (void)hipEventDestroy(sv->ev_upload);
    ctx->servers.erase(sv);
    delete sv;
    return PSF_OK;
}

int psf_server_add_basis(psf_server* sv, const psf_scene_desc* desc, const double* Phi, const double* lambda,
                         const double* gravity, const double* C, uint32_t* basis_id)
{
    if (!sv || !desc || !Phi || !lambda || !C || !basis_id) return fail(sv ? sv->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_ctx* ctx = sv->ctx;
    psf_scene_desc d = *desc;
    d.c_precision = PSF_C_BF16X2;
    d.phi_precision = PSF_PHI_FP32;
    uint64_t probe = 0;
    int rc = psf_scene_footprint(&d, &probe);   // validates the grid and K0
    if (rc) return fail(ctx, rc, "basis descriptor is invalid");
    if (d.K > 1024) return fail(ctx, PSF_ERR_UNSUPPORTED, "basis K0 must be at most 1024");
    if (sv->cfg.shared_basis && !group_supported(static_cast<int>(d.K)))
        return fail(ctx, PSF_ERR_UNSUPPORTED, "shared-basis steps need K0 <= 256");
    SRV_HIPCHK(ctx, hipSetDevice(ctx->device));
    Basis b;
    b.desc = d;
    b.F = (d.nx + 1) * d.ny * d.nz + (d.ny + 1) * d.nx * d.nz + (d.nz + 1) * d.nx * d.ny;
    b.ldF = (size_t(b.F) + 255) / 256 * 256;
    const size_t K0 = d.K;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b.Phi), b.ldF * K0 * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b.C), K0 * K0 * K0 * 8);
    if (e == hipSuccess) e = hipMemset(b.Phi, 0, b.ldF * K0 * 4);
    if (e == hipSuccess) {
        std::vector<float> col(b.F);
        for (size_t k = 0; k < K0 && e == hipSuccess; ++k) {
            for (uint32_t f = 0; f < b.F; ++f) col[f] = static_cast<float>(Phi[k * b.F + f]);
            e = hipMemcpy(b.Phi + k * b.ldF, col.data(), size_t(b.F) * 4, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess) e = hipMemcpy(b.C, C, K0 * K0 * K0 * 8, hipMemcpyHostToDevice);
    b.lambda.assign(lambda, lambda + K0);
    if (gravity)
        b.gravity.assign(gravity, gravity + K0);
    else
        b.gravity.assign(K0, 0.0);
    if (sv->cfg.shared_basis) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b.gravity_dev), K0 * 8);
        if (e == hipSuccess) e = hipMemcpy(b.gravity_dev, b.gravity.data(), K0 * 8, hipMemcpyHostToDevice);
    }
    std::string gerr;
    if (e == hipSuccess && sv->cfg.shared_basis) {
        rc = group_basis_create(b.C, static_cast<int>(K0), b.gravity_dev, ctx->stream, &b.group, &gerr);
        if (rc != PSF_OK && gerr.empty()) gerr = "shared-basis blocks: out of memory";
    }
    if (e != hipSuccess || rc != PSF_OK) {
        if (b.Phi) (void)hipFree(b.Phi);
        if (b.C) (void)hipFree(b.C);
        if (b.gravity_dev) (void)hipFree(b.gravity_dev);
        if (e != hipSuccess)
            return fail(ctx, e == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP,
                        std::string("basis upload: ") + hipGetErrorString(e));
        return fail(ctx, rc, gerr);
    }
    sv->bases.push_back(std::move(b));
    *basis_id = static_cast<uint32_t>(sv->bases.size() - 1);
    return PSF_OK;
}

int psf_server_request_footprint(psf_server* sv, uint32_t basis_id, uint32_t K, uint32_t n, uint64_t* bytes)
{
    if (!sv || !bytes) return PSF_ERR_ARG;
    if (basis_id >= sv->bases.size()) return fail(sv->ctx, PSF_ERR_ARG, "unknown basis");
    const Basis& b = sv->bases[basis_id];
    if (K == 0 || K > b.desc.K) return fail(sv->ctx, PSF_ERR_ARG, "K must be in [1, K0]");
    *bytes = request_bytes(sv, b, K, n);
    return PSF_OK;
}

int psf_server_submit(psf_server* sv, uint32_t basis_id, uint32_t K, uint32_t nframes, double dt, double viscosity,
                      int gravity_active, float lifetime, const double* omega, const float* particles, uint32_t n,
                      uint64_t* request_id)
{
    if (!sv || !omega || (!particles && n) || !request_id) return fail(sv ? sv->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_ctx* ctx = sv->ctx;
    if (basis_id >= sv->bases.size()) return fail(ctx, PSF_ERR_ARG, "unknown basis");
    const Basis& b = sv->bases[basis_id];
    if (K == 0 || K > b.desc.K) return fail(ctx, PSF_ERR_ARG, "K must be in [1, K0]");
    if (nframes == 0) return fail(ctx, PSF_ERR_ARG, "nframes must be positive");
    if (!std::isfinite(dt) || !std::isfinite(viscosity)) return fail(ctx, PSF_ERR_ARG, "non-finite parameter");
    const int ipw = request_ipw(sv, K);
    const int wgs = sv->cfg.shared_basis ? 0 : persistent_workgroups(PSF_C_BF16X2, static_cast<int>(K), ipw);
    const uint64_t fp = request_bytes(sv, b, K, n);
    if (wgs > sv->wg_capacity) return fail(ctx, PSF_ERR_UNSUPPORTED, "request needs more workgroups than one launch holds");
    if (fp > sv->cfg.mem_budget) return fail(ctx, PSF_ERR_NOMEM, "request footprint exceeds the server's memory budget");
    auto r = std::make_unique<Request>();
    r->id = sv->next_id++;
    r->basis = basis_id;
    r->K = K;
    r->nframes = nframes;
    r->n = n;
    r->dt = dt;
    r->nu = viscosity;
    r->grav = gravity_active ? 1 : 0;
    r->lifetime = lifetime;
    r->omega.assign(omega, omega + K);
    r->parts.assign(size_t(n) * 4, 0.0f);
    // storage order now, on the submitting thread: admission only copies
    r->sorted.resize(size_t(n) * 4);
    r->pid.resize(n);
    psf_scene_desc sd = b.desc;
    sort_particles(sd, particles, n, r->sorted.data(), r->pid.data());
    r->footprint = fp;
    r->wgs = wgs;
    r->ipw = ipw;
    r->t_submit = Clock::now();
    Request* raw = r.get();
    sv->reqs.emplace(raw->id, std::move(r));
    sv->queue.push_back(raw);
    sv->st.submitted++;
    *request_id = raw->id;
    return PSF_OK;
}

int psf_server_round(psf_server* sv, uint32_t* completed)
{
    if (!sv) return PSF_ERR_ARG;
    psf_ctx* ctx = sv->ctx;
    const Clock::time_point t0 = Clock::now();
    if (completed) *completed = 0;
    SRV_HIPCHK(ctx, hipSetDevice(ctx->device));
    // with nothing on the device, admit first so that this call launches work
    if (sv->active.empty()) {
        admit_queued(sv, true);
        sv->st.admit_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }

    // ---- launch the next round for every scene with frames left ----------------
    std::vector<Request*> run;
    for (Request* r : sv->active)
        if (r->sched < r->nframes) run.push_back(r);
    const int n = static_cast<int>(run.size());
    const bool launch = n > 0;
    if (launch) {
        const int W = static_cast<int>(sv->cfg.window);
        const int q = static_cast<int>(sv->launched & 1);
        hipStream_t sa = ctx->stream, sb = sv->second, sr = sv->third;
        std::vector<psf_scene*> sc(n);
        std::vector<int> nsteps(n), ld(n, W);
        std::vector<double*> hist(n);
        std::vector<float*> uh(n);
        uint32_t wblocks = 0, tblocks = 0;
        for (int j = 0; j < n; ++j) {
            Request* r = run[j];
            psf_scene* s = r->scene;
            sc[j] = s;
            nsteps[j] = static_cast<int>(std::min<uint32_t>(static_cast<uint32_t>(W), r->nframes - r->sched));
            hist[j] = q == 0 ? s->dev.hist : r->hist2;
            uh[j] = q == 0 ? s->dev.u_hist : r->u_hist2;
            TraceScene t{};
            t.P = scene_trace_params(s);
            t.u_hist = uh[j];
            t.ldU = s->dev.ldF;
            t.parts = s->dev.parts;
            t.pid = s->dev.pid;
            t.n = s->n_particles;
            t.nframes = static_cast<uint32_t>(nsteps[j]);
            t.block0 = tblocks;
            sv->h_ttab[q][j] = t;
            tblocks += (s->n_particles + 255) / 256;
        }
        // velocity fields: with FP32 sums the scenes of one basis go in pairs
        // (by K, largest first), each pair one pass over the basis's Phi
        int nw = 0;
        {
            std::vector<std::vector<int>> by_basis(sv->bases.size());
            for (int j = 0; j < n; ++j) by_basis[run[j]->basis].push_back(j);
            const size_t step = sv->cfg.u_fp32_sums ? 2 : 1;
            for (std::vector<int>& js : by_basis) {
                std::stable_sort(js.begin(), js.end(), [&](int a, int c) { return sc[a]->d.K > sc[c]->d.K; });
                for (size_t x = 0; x < js.size(); x += step) {
                    const int a = js[x];
                    psf_scene* s = sc[a];
                    WindowScene w{sv->bases[run[a]->basis].Phi, s->dev.ldF, s->F, static_cast<int>(s->d.K), hist[a],
                                  uh[a], wblocks, 0, nullptr, nullptr};
                    if (step == 2 && x + 1 < js.size()) {
                        const int c = js[x + 1];
                        w.K2 = static_cast<int>(sc[c]->d.K);
                        w.hist2 = hist[c];
                        w.uh2 = uh[c];
                    }
                    sv->h_wtab[q][nw++] = w;
                    wblocks += window_blocks(W, s->dev.ldF);
                }
            }
        }
        // the round two back used this parity's history buffers: P waits for
        // its reconstruction (omega history), R for its particles (u history)
        if (sv->launched >= 2) {
            SRV_HIPCHK(ctx, hipStreamWaitEvent(sa, sv->ev_rec[q], 0));
            SRV_HIPCHK(ctx, hipStreamWaitEvent(sr, sv->ev_done[q], 0));
        }
        if (sv->upload_pending) {   // newly admitted scenes' uploads
            SRV_HIPCHK(ctx, hipStreamWaitEvent(sa, sv->ev_upload, 0));
            sv->upload_pending = false;
        }
        if (sv->cfg.shared_basis) {
            // instances: each basis's scenes in runs of at most kGroupMaxScenes, as even as possible
            std::vector<std::vector<int>> by_basis(sv->bases.size());
            for (int j = 0; j < n; ++j) by_basis[run[j]->basis].push_back(j);
            std::vector<GroupScene> gs(n);
            std::vector<GroupInstance> inst;
            int o = 0;
            for (size_t bi = 0; bi < by_basis.size(); ++bi) {
                const std::vector<int>& js = by_basis[bi];
                const int m = static_cast<int>(js.size());
                if (!m) continue;
                const int ni = (m + kGroupMaxScenes - 1) / kGroupMaxScenes;
                for (int c = 0, x = 0; c < ni; ++c) {
                    const int cnt = m / ni + (c < m % ni ? 1 : 0);
                    inst.push_back(GroupInstance{sv->bases[bi].group, cnt, gs.data() + o});
                    for (int y = 0; y < cnt; ++y, ++x, ++o) {
                        const int j = js[x];
                        psf_scene* s = sc[j];
                        gs[o] = GroupScene{s->dev.omega, s->dev.visc, hist[j], s->dev.status, s->dt,
                                           static_cast<int>(s->d.K), nsteps[j], s->gravity_active, W};
                    }
                }
            }
            SRV_HIPCHK(ctx, launch_group_steps(sv->gbatch[q], inst.data(), static_cast<int>(inst.size()), wait_ticks(sv),
                                               sa));
        } else {
            SRV_HIPCHK(ctx, launch_persistent_batch(sv->batch[q], sc.data(), nsteps.data(), hist.data(), ld.data(), n, sa));
        }
        SRV_HIPCHK(ctx, hipEventRecord(sv->ev_step[q], sa));
        SRV_HIPCHK(ctx, hipStreamWaitEvent(sr, sv->ev_step[q], 0));
        SRV_HIPCHK(ctx, hipMemcpyAsync(sv->d_wtab[q], sv->h_wtab[q], sizeof(WindowScene) * nw, hipMemcpyHostToDevice, sr));
        SRV_HIPCHK(ctx, launch_reconstruct_window_multi(sv->cfg.u_fp32_sums ? PSF_PHI_FP32_ACC32 : PSF_PHI_FP32, W,
                                                        sv->d_wtab[q], nw, wblocks, sr));
        SRV_HIPCHK(ctx, hipEventRecord(sv->ev_rec[q], sr));
        SRV_HIPCHK(ctx, hipStreamWaitEvent(sb, sv->ev_rec[q], 0));
        SRV_HIPCHK(ctx, hipMemcpyAsync(sv->d_ttab[q], sv->h_ttab[q], sizeof(TraceScene) * n, hipMemcpyHostToDevice, sb));
        SRV_HIPCHK(ctx, launch_trace_window_multi(sv->d_ttab[q], n, tblocks, sb));
        for (int j = 0; j < n; ++j) {
            Request* r = run[j];
            r->scene->frame += static_cast<uint32_t>(nsteps[j]);
            r->sched += static_cast<uint32_t>(nsteps[j]);
            sv->st.scene_frames += static_cast<uint64_t>(nsteps[j]);
            if (r->sched == r->nframes) {
                r->last_round = sv->launched;
                sv->wgs_in_use -= r->wgs;
                // its results into its pinned buffer behind this round's particle
                // launch (omega and status are final since the round's steps):
                // retiring it then waits for nothing on the device
                psf_scene* s = r->scene;
                unsigned char* pin = static_cast<unsigned char*>(r->pin);
                const size_t np = size_t(r->n) * 16, nw = size_t(r->K) * 8;
                if (np) SRV_HIPCHK(ctx, hipMemcpyAsync(pin, s->dev.parts, np, hipMemcpyDeviceToHost, sb));
                SRV_HIPCHK(ctx, hipMemcpyAsync(pin + np, s->dev.omega, nw, hipMemcpyDeviceToHost, sb));
                SRV_HIPCHK(ctx, hipMemcpyAsync(pin + np + nw, s->dev.status, sizeof(int), hipMemcpyDeviceToHost, sb));
            }
        }
        SRV_HIPCHK(ctx, hipEventRecord(sv->ev_done[q], sb));   // after the particles and the copies
        sv->launched++;
        sv->st.rounds++;
    }

    // ---- admit while the device runs the round; admitted scenes join the next ------
    {
        const Clock::time_point ta = Clock::now();
        admit_queued(sv, !launch);
        sv->st.admit_ms += std::chrono::duration<double, std::milli>(Clock::now() - ta).count();
    }

    // ---- hand out the requests of the round before (all rounds when idle) ----------
    const int64_t target = launch ? sv->launched - 2 : sv->launched - 1;
    if (target > sv->retired) {
        SRV_HIPCHK(ctx, hipEventSynchronize(sv->ev_done[target & 1]));
        const Clock::time_point t2 = Clock::now();
        std::vector<Request*> keep;
        keep.reserve(sv->active.size());
        for (Request* r : sv->active) {
            if (r->last_round < 0 || r->last_round > target) {
                keep.push_back(r);
                continue;
            }
            const int rc = read_results(sv, r);
            r->t_done = Clock::now();
            sv->finished.push_back(r->id);
            if (rc == PSF_OK) {
                r->state = kDone;
                sv->st.completed++;
                if (completed) (*completed)++;
            } else {
                r->state = kFailed;
                r->error = ctx->err;
                if (sv->cfg.shared_basis) {
                    // what the launch's first wait that gave up waited for
                    for (int q2 = 0; q2 < 2; ++q2) {
                        const std::string d = group_batch_diag(sv->gbatch[q2]);
                        if (!d.empty()) sv->last_diag = d;
                    }
                    if (!sv->last_diag.empty()) r->error += " (" + sv->last_diag + ")";
                }
                sv->st.failed++;
            }
            release(sv, r);
        }
        sv->active.swap(keep);
        sv->retired = target;
        sv->st.retire_ms += std::chrono::duration<double, std::milli>(Clock::now() - t2).count();
    }
    sv->st.busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    return PSF_OK;
}

int psf_server_done(psf_server* sv, uint64_t id)
{
    if (!sv) return PSF_ERR_ARG;
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(sv->ctx, PSF_ERR_ARG, "unknown request");
    return it->second->state == kDone || it->second->state == kFailed ? 1 : 0;
}

int psf_server_result(psf_server* sv, uint64_t id, double* omega, float* particles, double* latency_ms)
{
    if (!sv) return PSF_ERR_ARG;
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(sv->ctx, PSF_ERR_ARG, "unknown request");
    Request* r = it->second.get();
    if (r->state == kQueued || r->state == kActive) return fail(sv->ctx, PSF_ERR_STATE, "request has not finished");
    if (r->state == kFailed) {
        const std::string msg = "request failed: " + r->error;
        sv->reqs.erase(it);
        return fail(sv->ctx, PSF_ERR_HIP, msg);
    }
    if (omega) std::memcpy(omega, r->omega.data(), r->omega.size() * 8);
    if (particles) std::memcpy(particles, r->parts.data(), r->parts.size() * 4);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    sv->reqs.erase(it);
    return PSF_OK;
}

int psf_server_completed(psf_server* sv, uint64_t* ids, uint32_t capacity, uint32_t* count)
{
    if (!sv || !count || (!ids && capacity)) return PSF_ERR_ARG;
    uint32_t c = 0;
    while (c < capacity && !sv->finished.empty()) {
        ids[c++] = sv->finished.front();
        sv->finished.pop_front();
    }
    *count = c;
    return PSF_OK;
}

int psf_server_get_stats(psf_server* sv, psf_server_stats* out)
{
    if (!sv || !out) return PSF_ERR_ARG;
    psf_server_stats s = sv->st;
    s.active = static_cast<uint32_t>(sv->active.size());
    s.queued = static_cast<uint32_t>(sv->queue.size());
    s.mem_in_use = sv->mem_in_use;
    s.mem_peak = sv->mem_peak;
    *out = s;
    return PSF_OK;
}

}  // extern "C"
