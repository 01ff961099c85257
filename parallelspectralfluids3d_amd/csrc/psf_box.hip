// Analytic closed-box basis and its structure tensor, built on the device
// (C ABI in psf.h).
//
// The reference obtains its basis from an eigensolve of the DEC 1-form
// Laplacian (psfsolver.cpp:236-342) and C from a voxel sum
// (psfsolver.cpp:431-551).  For a closed box the discrete sin/cos modes are a
// divergence-free basis with closed forms, and every quantity in C is a
// product of 1-D profiles along x, y and z:
//   - a mode's flux on a face normal to axis A is amp_A * (node sine along A)
//     * (cell cosine along the other two axes);
//   - V = curl Phi on an edge is Hf times four signed face fluxes at fixed
//     offsets (kCurlFaces), negated on edges that lie on a low wall plane;
//   - each of the 12 voxel slots (edge e_s, faces a_s, b_s) contributes
//     sum_v V[e_s(v), k] Phi[a_s(v), i] Phi[b_s(v), j], which factorises into a
//     product over the three axes of 1-D triple sums S[f_k, f_i, f_j] over the
//     distinct frequencies of that axis.
// So Phi is written straight from 1-D tables, and C_i(k, j) is a sum over 48
// (slot, curl term) pairs of table products, one thread per entry.  Nothing
// of size F or E is formed on the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "psf_internal.hpp"
#include "psf_kernels.hpp"
#include "psf_persistent.hpp"

namespace psf {
namespace {

constexpr double kPi = 3.141592653589793;

// face slots f1..f6 of voxel (x,y,z): (axis 0=x 1=y 2=z, offset along it)
constexpr int kFaceAxis[6] = {2, 2, 1, 1, 0, 0};
constexpr int kFaceSide[6] = {0, 1, 0, 1, 0, 1};
// the 12 voxel slots: (face holding the edge, edge slot, face a, face b) -- advection.SLOTS
constexpr int kSlotFaces[12][2] = {{0, 3}, {5, 0}, {2, 0}, {0, 4}, {3, 1}, {4, 1},
                                   {1, 2}, {1, 5}, {4, 3}, {2, 4}, {3, 5}, {5, 2}};
// edge of each slot for interior creators: (edge axis, dx, dy, dz) relative to the voxel
constexpr int kSlotEdge[12][4] = {{0, 0, 1, 0}, {1, 0, 0, 0}, {0, 0, 0, 0}, {1, 1, 0, 0},
                                  {0, 0, 1, 1}, {1, 0, 0, 1}, {0, 0, 0, 1}, {1, 1, 0, 1},
                                  {2, 0, 1, 0}, {2, 0, 0, 0}, {2, 1, 1, 0}, {2, 1, 0, 0}};
// stored normals of interior faces (BACK +z, TOP -y, RIGHT -x), by face axis
constexpr double kInteriorNormal[3][3] = {{-1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}, {0.0, 0.0, 1.0}};
// faces around an edge of axis A: (face axis, face coordinates - edge coordinates, sign)
struct CurlFace {
    int axis, d[3];
    double sign;
};
constexpr CurlFace kCurlFaces[3][4] = {
    {{1, {0, 0, -1}, -1.0}, {1, {0, 0, 0}, 1.0}, {2, {0, -1, 0}, 1.0}, {2, {0, 0, 0}, -1.0}},
    {{0, {0, 0, -1}, -1.0}, {0, {0, 0, 0}, 1.0}, {2, {-1, 0, 0}, 1.0}, {2, {0, 0, 0}, -1.0}},
    {{0, {0, -1, 0}, -1.0}, {0, {0, 0, 0}, 1.0}, {1, {-1, 0, 0}, 1.0}, {1, {0, 0, 0}, -1.0}}};

#pragma clang fp contract(off)
// node sine at node idx (zero at and beyond the walls), cell cosine at cell idx (zero outside)
double node_sine(int n, int f, int idx) { return (idx <= 0 || idx >= n) ? 0.0 : std::sin(f * kPi * idx / n); }
double cell_cos(int n, int f, int idx) { return (idx < 0 || idx > n - 1) ? 0.0 : std::cos(f * kPi * (idx + 0.5) / n); }

struct BoxModes {
    int dims[3];
    int K;
    std::vector<int> uniq[3];     // distinct frequencies per axis, ascending
    std::vector<int> pos;         // K x 3: index of each mode's frequency in uniq
};

int prepare(psf_scene* s, const int32_t* freqs, BoxModes& m)
{
    m.dims[0] = static_cast<int>(s->d.nx);
    m.dims[1] = static_cast<int>(s->d.ny);
    m.dims[2] = static_cast<int>(s->d.nz);
    m.K = static_cast<int>(s->d.K);
    m.pos.assign(size_t(m.K) * 3, 0);
    for (int a = 0; a < 3; ++a) {
        std::vector<int> f(m.K);
        for (int k = 0; k < m.K; ++k) {
            f[k] = freqs[3 * k + a];
            if (f[k] < 0 || f[k] >= m.dims[a]) return -1;
        }
        std::vector<int> u = f;
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        m.uniq[a] = u;
        for (int k = 0; k < m.K; ++k)
            m.pos[3 * k + a] = static_cast<int>(std::lower_bound(u.begin(), u.end(), f[k]) - u.begin());
    }
    return 0;
}

// one (slot, curl term) pair of the separable sum
struct BoxTerm {
    double sgn_hf;   // curl sign * Hf
    double shc;      // s_h * c_s
    int face, fa, fb;   // face axes of the curl term, of face a and of face b
    int off[3];      // offsets of the three axis tables in the table array
};

void build_terms(const BoxModes& m, double Hf, double sh, std::vector<BoxTerm>& terms, std::vector<double>& tables)
{
    for (int s = 0; s < 12; ++s) {
        const int a = kSlotFaces[s][0], b = kSlotFaces[s][1];
        const int axa = kFaceAxis[a], axb = kFaceAxis[b];
        const double* na = kInteriorNormal[axa];
        const double* nb = kInteriorNormal[axb];
        const double cr[3] = {na[1] * nb[2] - na[2] * nb[1], na[2] * nb[0] - na[0] * nb[2], na[0] * nb[1] - na[1] * nb[0]};
        const double c_s = (cr[0] + cr[1]) + cr[2];
        if (c_s == 0.0) continue;
        const int ea = kSlotEdge[s][0];
        const int off[3] = {kSlotEdge[s][1], kSlotEdge[s][2], kSlotEdge[s][3]};
        for (int t = 0; t < 4; ++t) {
            const CurlFace& cf = kCurlFaces[ea][t];
            BoxTerm T{};
            T.sgn_hf = cf.sign * Hf;
            T.shc = sh * c_s;
            T.face = cf.axis;
            T.fa = axa;
            T.fb = axb;
            for (int al = 0; al < 3; ++al) {
                const int n = m.dims[al];
                const auto& U = m.uniq[al];
                const int D = static_cast<int>(U.size());
                T.off[al] = static_cast<int>(tables.size());
                std::vector<double> Q(size_t(D) * n), Pa(size_t(D) * n), Pb(size_t(D) * n);
                for (int r = 0; r < D; ++r)
                    for (int v = 0; v < n; ++v) {
                        const int ec = v + off[al], fc = ec + cf.d[al];
                        double q = al == cf.axis ? node_sine(n, U[r], fc) : cell_cos(n, U[r], fc);
                        if (al != ea && ec == 0) q = -q;   // edge on a low wall plane
                        Q[size_t(r) * n + v] = q;
                        Pa[size_t(r) * n + v] = al == axa ? node_sine(n, U[r], v + kFaceSide[a]) : cell_cos(n, U[r], v);
                        Pb[size_t(r) * n + v] = al == axb ? node_sine(n, U[r], v + kFaceSide[b]) : cell_cos(n, U[r], v);
                    }
                for (int fk = 0; fk < D; ++fk)
                    for (int fi = 0; fi < D; ++fi)
                        for (int fj = 0; fj < D; ++fj) {
                            double acc = 0.0;
                            for (int v = 0; v < n; ++v)
                                acc += Q[size_t(fk) * n + v] * Pa[size_t(fi) * n + v] * Pb[size_t(fj) * n + v];
                            tables.push_back(acc);
                        }
            }
            terms.push_back(T);
        }
    }
}

// C[(i*K + j)*K + k] = C_i(k, j): one thread per entry, k fastest
__global__ __launch_bounds__(256) void k_box_advection(const BoxTerm* __restrict__ terms, int nterms,
                                                       const double* __restrict__ tab, int Dx, int Dy, int Dz,
                                                       const int* __restrict__ pos, const double* __restrict__ amp,
                                                       const double* __restrict__ lambda, int K, double* __restrict__ C)
{
#pragma clang fp contract(off)
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const size_t KK = static_cast<size_t>(K) * K;
    if (t >= KK * K) return;
    const int k = static_cast<int>(t % K), j = static_cast<int>((t / K) % K), i = static_cast<int>(t / KK);
    const int pk[3] = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
    const int pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const int pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    const int D[3] = {Dx, Dy, Dz};
    double acc = 0.0;
    for (int q = 0; q < nterms; ++q) {
        const BoxTerm T = terms[q];
        double pij = 1.0, pji = 1.0;
#pragma unroll
        for (int al = 0; al < 3; ++al) {
            const double* S = tab + T.off[al];
            pij = pij * S[(pk[al] * D[al] + pi[al]) * D[al] + pj[al]];
            pji = pji * S[(pk[al] * D[al] + pj[al]) * D[al] + pi[al]];
        }
        const double ak = T.sgn_hf * amp[3 * k + T.face];
        const double tij = ((ak * amp[3 * i + T.fa]) * amp[3 * j + T.fb]) * pij;
        const double tji = ((ak * amp[3 * j + T.fa]) * amp[3 * i + T.fb]) * pji;
        acc = acc + (tij - tji) * T.shc;
    }
    C[t] = acc / lambda[i];
}

// Phi[k*ldF + f] from the 1-D profile tables: node sines [axis][D][n+1] at
// prof + noff[axis], cell cosines [axis][D][n] at prof + coff[axis]
template <typename T>
__global__ __launch_bounds__(256) void k_box_basis(T* __restrict__ Phi, size_t ldF, uint32_t F, uint32_t nZF,
                                                   uint32_t nYF, int nx, int ny, int nz, int K,
                                                   const int* __restrict__ pos, const double* __restrict__ amp,
                                                   const double* __restrict__ prof, int nx_off, int ny_off,
                                                   int nz_off, int cx_off, int cy_off, int cz_off)
{
#pragma clang fp contract(off)
    const uint32_t f = blockIdx.x * 256u + threadIdx.x;
    if (f >= F) return;
    int axis, x, y, z;
    if (f < nZF) {
        axis = 2;
        z = static_cast<int>(f / (nx * ny));
        const int r = static_cast<int>(f % (nx * ny));
        y = r / nx;
        x = r % nx;
    } else if (f < nZF + nYF) {
        axis = 1;
        const uint32_t g = f - nZF;
        z = static_cast<int>(g / ((ny + 1) * nx));
        const int r = static_cast<int>(g % ((ny + 1) * nx));
        y = r / nx;
        x = r % nx;
    } else {
        axis = 0;
        const uint32_t g = f - nZF - nYF;
        z = static_cast<int>(g / (ny * (nx + 1)));
        const int r = static_cast<int>(g % (ny * (nx + 1)));
        y = r / (nx + 1);
        x = r % (nx + 1);
    }
    for (int k = 0; k < K; ++k) {
        const int fx = pos[3 * k], fy = pos[3 * k + 1], fz = pos[3 * k + 2];
        double v;
        if (axis == 2) {          // Z faces: node sine in z, cell cosines in y and x
            v = amp[3 * k + 2] * ((prof[nz_off + fz * (nz + 1) + z] * prof[cy_off + fy * ny + y]) * prof[cx_off + fx * nx + x]);
        } else if (axis == 1) {   // Y faces
            v = amp[3 * k + 1] * ((prof[cz_off + fz * nz + z] * prof[ny_off + fy * (ny + 1) + y]) * prof[cx_off + fx * nx + x]);
        } else {                  // X faces
            v = amp[3 * k] * ((prof[cz_off + fz * nz + z] * prof[cy_off + fy * ny + y]) * prof[nx_off + fx * (nx + 1) + x]);
        }
        Phi[static_cast<size_t>(k) * ldF + f] = static_cast<T>(v);
    }
}

__global__ __launch_bounds__(256) void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t < n) b[t] = static_cast<float>(a[t]);
}

int fail(psf_ctx* ctx, int code, const std::string& msg) { return ctx_set_error(ctx, code, msg); }

#define BOX_HIPCHK(ctx, expr)                                                                                 \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return fail((ctx), PSF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// small device copies that live for one call
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

template <typename V>
hipError_t upload(DevBuf& b, const std::vector<V>& v)
{
    hipError_t e = hipMalloc(&b.p, std::max<size_t>(v.size(), 1) * sizeof(V));
    if (e == hipSuccess && !v.empty()) e = hipMemcpy(b.p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice);
    return e;
}

}  // namespace
}  // namespace psf

using namespace psf;

extern "C" {

int psf_scene_generate_box_basis(psf_scene* s, const int32_t* freqs, const double* amps, const double* lambda,
                                 const double* gravity)
{
    if (!s || !freqs || !amps || !lambda) return fail(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_ctx* ctx = s->ctx;
    if (s->phi_borrowed) return fail(ctx, PSF_ERR_STATE, "the scene's Phi belongs to a server basis");
    BoxModes m;
    if (prepare(s, freqs, m)) return fail(ctx, PSF_ERR_ARG, "a mode frequency lies outside [0, n) of its axis");
    BOX_HIPCHK(ctx, hipSetDevice(ctx->device));
    // Phi changes: frames computed ahead are dropped (and the side stream's
    // reads of Phi come before the writes below)
    if (int rc = scene_window_flush(s)) return rc;
    // 1-D profiles: node sines then cell cosines of each axis
    std::vector<double> prof;
    int noff[3], coff[3];
    for (int a = 0; a < 3; ++a) {
        const int n = m.dims[a];
        noff[a] = static_cast<int>(prof.size());
        for (int f : m.uniq[a])
            for (int i = 0; i <= n; ++i) prof.push_back(node_sine(n, f, i));
        coff[a] = static_cast<int>(prof.size());
        for (int f : m.uniq[a])
            for (int i = 0; i < n; ++i) prof.push_back(cell_cos(n, f, i));
    }
    std::vector<double> amp(amps, amps + size_t(m.K) * 3);
    DevBuf dpos, damp, dprof;
    BOX_HIPCHK(ctx, upload(dpos, m.pos));
    BOX_HIPCHK(ctx, upload(damp, amp));
    BOX_HIPCHK(ctx, upload(dprof, prof));
    const unsigned grid = (s->F + 255) / 256;
    if (s->d.phi_precision != PSF_PHI_FP64)   // FP32 storage
        k_box_basis<float><<<grid, 256, 0, ctx->stream>>>(
            static_cast<float*>(s->dev.Phi), s->dev.ldF, s->F, s->nZF, s->nYF, m.dims[0], m.dims[1], m.dims[2], m.K,
            static_cast<const int*>(dpos.p), static_cast<const double*>(damp.p), static_cast<const double*>(dprof.p),
            noff[0], noff[1], noff[2], coff[0], coff[1], coff[2]);
    else
        k_box_basis<double><<<grid, 256, 0, ctx->stream>>>(
            static_cast<double*>(s->dev.Phi), s->dev.ldF, s->F, s->nZF, s->nYF, m.dims[0], m.dims[1], m.dims[2], m.K,
            static_cast<const int*>(dpos.p), static_cast<const double*>(damp.p), static_cast<const double*>(dprof.p),
            noff[0], noff[1], noff[2], coff[0], coff[1], coff[2]);
    BOX_HIPCHK(ctx, hipGetLastError());
    BOX_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    s->has_phi = true;
    // lambda, gravity and the viscosity factors as for an uploaded basis
    return psf_scene_upload_basis(s, nullptr, lambda, gravity);
}

int psf_scene_build_box_advection(psf_scene* s, const int32_t* freqs, const double* amps, const double* lambda)
{
    if (!s || !freqs || !amps || !lambda) return fail(s ? s->ctx : nullptr, PSF_ERR_ARG, "null argument");
    psf_ctx* ctx = s->ctx;
    const int prec = s->d.c_precision;
    if (prec == PSF_C_FP16) return fail(ctx, PSF_ERR_UNSUPPORTED, "FP16 C needs per-mode scales: upload it instead");
    BoxModes m;
    if (prepare(s, freqs, m)) return fail(ctx, PSF_ERR_ARG, "a mode frequency lies outside [0, n) of its axis");
    for (int k = 0; k < m.K; ++k)
        if (!(lambda[k] != 0.0) || !std::isfinite(lambda[k])) return fail(ctx, PSF_ERR_ARG, "lambda must be finite and non-zero");
    BOX_HIPCHK(ctx, hipSetDevice(ctx->device));
    if (int rc = scene_window_flush(s)) return rc;   // C changes: as in psf_scene_generate_box_basis
    // Hf = 1/(2h) and s_h = (h/2)^2 with h the float resolution (dec.cpp:143-193, psfsolver.cpp:444)
    const float hf32 = static_cast<float>(s->d.resolution);
    const double Hf = 1.0 / (2.0 * static_cast<double>(hf32));
    const float hr = hf32 / 2.0f;
    const double sh = static_cast<double>(hr * hr);
    std::vector<BoxTerm> terms;
    std::vector<double> tables;
    build_terms(m, Hf, sh, terms, tables);
    std::vector<double> amp(amps, amps + size_t(m.K) * 3), lam(lambda, lambda + m.K);
    DevBuf dterms, dtab, dpos, damp, dlam, dC;
    BOX_HIPCHK(ctx, upload(dterms, terms));
    BOX_HIPCHK(ctx, upload(dtab, tables));
    BOX_HIPCHK(ctx, upload(dpos, m.pos));
    BOX_HIPCHK(ctx, upload(damp, amp));
    BOX_HIPCHK(ctx, upload(dlam, lam));
    const size_t n = size_t(m.K) * m.K * m.K;
    double* Cf = nullptr;
    if (prec == PSF_C_FP64) {
        Cf = static_cast<double*>(s->dev.C);
    } else {
        hipError_t e = hipMalloc(&dC.p, n * sizeof(double));
        if (e != hipSuccess)
            return fail(ctx, e == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP,
                        std::string("temporary C: ") + hipGetErrorString(e));
        Cf = static_cast<double*>(dC.p);
    }
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    k_box_advection<<<grid, 256, 0, ctx->stream>>>(static_cast<const BoxTerm*>(dterms.p), static_cast<int>(terms.size()),
                                                   static_cast<const double*>(dtab.p),
                                                   static_cast<int>(m.uniq[0].size()), static_cast<int>(m.uniq[1].size()),
                                                   static_cast<int>(m.uniq[2].size()), static_cast<const int*>(dpos.p),
                                                   static_cast<const double*>(damp.p), static_cast<const double*>(dlam.p),
                                                   m.K, Cf);
    BOX_HIPCHK(ctx, hipGetLastError());
    if (prec == PSF_C_FP32) {
        k_f64_to_f32<<<grid, 256, 0, ctx->stream>>>(Cf, static_cast<float*>(s->dev.C), n);
        BOX_HIPCHK(ctx, hipGetLastError());
    } else if (prec != PSF_C_FP64) {
        BOX_HIPCHK(ctx, launch_pack_c(s, Cf, m.K, ctx->stream));
    }
    BOX_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    s->has_C = true;
    return PSF_OK;
}

}  // extern "C"
