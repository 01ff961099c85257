// Basis-shared persistent reduced step: the scene server's batched path.
//
// A server's scenes on one basis use its first K <= K0 modes, and their C_i(a, b)
// is the basis's C_i(a, b) for i, a, b < K (psf_server_add_basis).  With
// w_k = 0 for k >= K the K0-mode contraction gives exactly the scene's y_i for
// i < K: every term that holds a zero coefficient is an exact zero.  So one
// register-resident copy of the basis's C blocks (BF16 hi+lo, the layout of
// psf_persistent.hip) serves several scenes at once.  The MFMA's 16 B columns,
// of which the single-scene step fills two (w_hi, w_lo), carry the w_hi and
// w_lo of eight scenes: scene s is column pair s mod 8 of column set s / 8.
//
// An instance is one basis with up to kGroupMaxScenes scenes on G workgroups;
// a launch holds several instances (wg_map[blockIdx.x] = instance << 16 | g),
// each with its own exchange records and tags.  Workgroup g owns modes
// g, g + G, ... of the basis (psfsolver.cpp:39-44 for each scene):
//   1. contraction: each wave runs its register-resident blocks against every
//      column set; at the end of a row tile each lane contracts its
//      accumulator rows with its own scene's w in FP64, and at the end of a
//      mode the lanes of each scene are summed;
//   2. publish: y_i of every active scene, as two tagged 64-bit halves (the
//      single-scene step's record format, 16 bytes per scene and mode);
//   3. gather: every wave reads whole records (two per wave-wide load,
//      coalesced) until their tags match and leaves y in LDS;
//   4. update: wave s mod 8 applies the FP64 update of psfsolver.cpp:34-59
//      (energy, Euler, renormalisation, viscosity, forcing) to the
//      workgroup's copy of scene s's w.
// A scene that has run its steps, or whose energy became non-finite, keeps
// its w frozen for the rest of the launch but stays in the exchange.
//
// Every launched workgroup must be waited on by every other one at every
// step: the records alternate between two buffers by step parity, so a
// workgroup may only run one step ahead of the slowest.  Pollers wait for
// modes k < K of every scene of the instance at every step, and only the
// workgroups owning a mode below the instance's largest K are launched
// (workgroup g owns mode g), so each launched workgroup publishes a record
// that all others wait for.  (A workgroup owning only modes no scene uses
// would not hold the others back: started late, behind the others' first
// steps, it would find its records overwritten.)
//
// A scene's arithmetic does not depend on its column, on the instance's other
// scenes or on their number: an MFMA output column depends only on its own B
// column, each reduction over lanes pairs the same lanes for every scene,
// and a scene's update runs in one wave in a fixed order.  A request
// therefore gives the same bits alone and in a mix.
//
// Residency: every workgroup of a launch must be resident at once (the
// exchange waits on all of them); the caller keeps the launch within
// group_batch_capacity().  Waits are bounded (2 s by default) and report a
// timed-out exchange in the scenes' status words.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "psf_group.hpp"
#include "psf_persistent.hpp"
#include "psf_wave.hpp"

namespace psf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kTile = 16;
constexpr int kNW = 8;                     // waves per workgroup: two per SIMD
constexpr int kWBS = kGroupMaxK + 8;       // B-table row stride (16-byte shift per column)
constexpr int kSWS = kGroupMaxK + 4;       // w row stride (32-byte shift per scene)
constexpr int kGMaxIpw = 8;                // modes per workgroup
constexpr int kGRB = 36;                   // largest register budget, in MFMA blocks per wave
constexpr int kGMeta = kNW * kGRB;         // every block of a workgroup is register-resident
constexpr int kXW = 2 * kGroupMaxScenes;   // 64-bit words per mode record
constexpr uint64_t kGWait = 200000000ull;  // s_memrealtime runs at 100 MHz: 2 s
constexpr int kDiagWords = 8;
constexpr int kPollQ = kGroupMaxK / (4 * kNW);   // four-record groups each wave gathers
constexpr int kSc1 = 16;                          // buffer cache policy sc1: agent coherence

}  // namespace

struct GroupArgs {
    const uint4* blocks;          // [G][bpw][64]
    const uint32_t* meta;         // [bpw]
    const double* gravity;        // [K0]
    unsigned long long* xbuf;     // [2][K0][kXW]
    const GroupScene* scenes;     // [S]
    unsigned long long* diag;     // the first wait that gave up in the launch (group_batch_diag)
    int S, K0, G, ipw, bpw, nsteps, instance;
    uint32_t tag0;
    uint64_t wait_ticks;
};

struct GroupBasis {
    int K0 = 0, G = 0, ipw = 0, bpw = 0;
    uint4* blocks = nullptr;
    uint32_t* meta = nullptr;
    double* inv_scale = nullptr;  // written by the packer (ones); unused by the step
    const double* gravity = nullptr;
};

struct GroupBatch {
    int max_instances = 0, max_wgs = 0, capacity = 0;
    GroupArgs* d_args = nullptr;
    GroupScene* d_scenes = nullptr;
    uint32_t* d_map = nullptr;
    GroupArgs* h_args = nullptr;     // pinned staging
    GroupScene* h_scenes = nullptr;
    uint32_t* h_map = nullptr;
    unsigned long long* xbuf = nullptr;   // max_instances slots of 2 * kGroupMaxK * kXW words
    unsigned long long* diag = nullptr;   // kDiagWords: the first exchange wait that gave up
    std::vector<uint32_t> tags;           // last tag used by each slot
};

// Sum over the lanes of one scene: its two columns (hi, lo) and the four row
// groups of the 16 x 16 accumulator.  Lane 2 (s mod 8) of row group 0 ends
// with the total of scene s of the column set; every scene's lanes are paired
// in the same order.
__device__ __forceinline__ double scene_total(double v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

template <int RB, int NCS>
__global__ __launch_bounds__(64 * kNW, 1) void k_group_step(const GroupArgs* __restrict__ tab,
                                                            const uint32_t* __restrict__ wg_map)
{
    constexpr int SMAX = 8 * NCS;
    constexpr int NT = 64 * kNW;
    constexpr int SPW = (SMAX + kNW - 1) / kNW;   // scenes gathered and updated by each wave
    constexpr int JK = kGroupMaxK / 64;           // modes per lane in the update
    constexpr int BC = NCS == 1 ? 6 : 4;          // blocks whose B operands are loaded ahead

    __shared__ double s_w[SMAX * kSWS];
    __shared__ double s_y[SMAX * kSWS];      // y of the step, gathered from the records
    __shared__ __align__(16) uint16_t s_wb[NCS * 16 * kWBS];
    __shared__ uint32_t s_meta[kGMeta];
    __shared__ double s_ypart[kNW][kGMaxIpw][SMAX];
    __shared__ int s_K[SMAX], s_n[SMAX], s_dead[SMAX];
    __shared__ int s_flag;

    const uint32_t map = wg_map[blockIdx.x];
    const GroupArgs a = tab[map >> 16];
    const int g = static_cast<int>(map & 0xFFFFu);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kg = lane >> 4;
    const int K0 = a.K0, S = a.S;
    const int Kp = (K0 + kTile - 1) / kTile * kTile;
    // this instance's exchange slot (2 kGroupMaxK records of kXW words) as a buffer
    const __amdgpu_buffer_rsrc_t xrec = __builtin_amdgcn_make_buffer_rsrc(
        a.xbuf, 0, static_cast<int>(sizeof(unsigned long long) * 2 * kGroupMaxK * kXW), 0x00020000);

    for (int b = tid; b < a.bpw; b += NT) s_meta[b] = a.meta[b];
    if (tid < SMAX) {
        s_K[tid] = tid < S ? a.scenes[tid].K : 0;
        s_n[tid] = tid < S ? a.scenes[tid].nsteps : 0;
        s_dead[tid] = 0;
    }
    if (tid == 0) s_flag = 0;
    for (int e = tid; e < kNW * kGMaxIpw * SMAX; e += NT) (&s_ypart[0][0][0])[e] = 0.0;

    // w of scene s into its row of s_w and its two columns of the B table
    auto split_one = [&](int s, int k, double w) {
        const uint16_t hi = f32_to_bf16_rne(static_cast<float>(w));
        const uint16_t lo = f32_to_bf16_rne(static_cast<float>(w - static_cast<double>(bf16_to_f32(hi))));
        uint16_t* row = s_wb + (s >> 3) * 16 * kWBS + 2 * (s & 7) * kWBS;
        row[k] = hi;
        row[kWBS + k] = lo;
    };
    __syncthreads();
    for (int e = tid; e < SMAX * Kp; e += NT) {
        const int s = e / Kp, k = e - s * Kp;
        const double w = k < s_K[s] ? a.scenes[s].omega[k] : 0.0;
        s_w[s * kSWS + k] = w;
        split_one(s, k, w);
    }

    // this wave's block range; slots past its last block hold zero blocks
    const int per = (a.bpw + kNW - 1) / kNW;
    const int b0 = wv * per;
    const int nblk = max(0, min(a.bpw - b0, per));
    const uint4* gblk = a.blocks + (static_cast<size_t>(g) * a.bpw + b0) * 64 + lane;
    uint4 areg[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) {
        areg[j] = uint4{0u, 0u, 0u, 0u};
        if (j < nblk) areg[j] = gblk[static_cast<size_t>(j) * 64];
    }
    __syncthreads();
    // fmask bit j: block j ends a row tile (flush), tmask bit j: it ends a mode
    const uint32_t mreg = lane < nblk ? s_meta[b0 + lane] : 0u;
    uint64_t fmask = 0, tmask = 0;
    for (int j = 0; j < nblk; ++j) {
        const int kj = __builtin_amdgcn_readlane(static_cast<int>(mreg), j) >> 16;
        const int kn = j + 1 < nblk ? (__builtin_amdgcn_readlane(static_cast<int>(mreg), j + 1) >> 16) : -1;
        if (kn != kj) fmask |= 1ull << j;
        if (kn < 0 || (kn >> 8) != (kj >> 8)) tmask |= 1ull << j;
    }
    // B operand of block meta m for column set c: lane (col, kg) reads 8
    // entries of tile q from its column's row (hi+lo blocks have q1 == q2)
    const int bsel = col * kWBS + (kg & 1) * 8;
    auto b_operand = [&](uint32_t m, int c) {
        const int q = static_cast<int>((m >> 8) & 255u);
        return *reinterpret_cast<const uint4*>(s_wb + c * 16 * kWBS + bsel + q * kTile);
    };
    auto row_contract = [&](const f32x4& acc, int p, int c) {
        const double* wr = s_w + (8 * c + (col >> 1)) * kSWS + p * kTile + 4 * kg;
        return static_cast<double>(acc[0]) * wr[0] + static_cast<double>(acc[1]) * wr[1] +
               static_cast<double>(acc[2]) * wr[2] + static_cast<double>(acc[3]) * wr[3];
    };

    for (int step = 0; step < a.nsteps; ++step) {
        const uint32_t tag = a.tag0 + static_cast<uint32_t>(step) + 1u;
        // ---- contraction of this wave's blocks against every column set --------
        {
            const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
            f32x4 acc[NCS][2];
            double part[NCS];
#pragma unroll
            for (int c = 0; c < NCS; ++c) {
                acc[c][0] = zero4;
                acc[c][1] = zero4;
                part[c] = 0.0;
            }
            uint4 bq[2][BC][NCS];
#pragma unroll
            for (int jj = 0; jj < BC && jj < RB; ++jj)
#pragma unroll
                for (int c = 0; c < NCS; ++c)
                    bq[0][jj][c] = b_operand(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), jj)), c);
#pragma unroll
            for (int c0 = 0; c0 < RB; c0 += BC) {
                const int cb = (c0 / BC) & 1;
#pragma unroll
                for (int jj = 0; jj < BC; ++jj)
                    if (c0 + BC + jj < RB)
#pragma unroll
                        for (int c = 0; c < NCS; ++c)
                            bq[cb ^ 1][jj][c] = b_operand(
                                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), c0 + BC + jj)), c);
#pragma unroll
                for (int jj = 0; jj < BC; ++jj) {
                    const int j = c0 + jj;
                    if (j >= RB) break;
#pragma unroll
                    for (int c = 0; c < NCS; ++c)
                        acc[c][j & 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, areg[j]), __builtin_bit_cast(bf16x8, bq[cb][jj][c]), acc[c][j & 1],
                            0, 0, 0);
                    if ((fmask >> j) & 1ull) {
                        const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), j));
                        const int p = static_cast<int>((m >> 16) & 255u);
#pragma unroll
                        for (int c = 0; c < NCS; ++c) {
                            part[c] += row_contract(acc[c][0] + acc[c][1], p, c);
                            acc[c][0] = zero4;
                            acc[c][1] = zero4;
                        }
                        if ((tmask >> j) & 1ull) {
                            const int t = static_cast<int>(m >> 24);
#pragma unroll
                            for (int c = 0; c < NCS; ++c) {
                                const double v = scene_total(part[c]);
                                if (kg == 0 && !(col & 1)) s_ypart[wv][t][8 * c + (col >> 1)] += v;
                                part[c] = 0.0;
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- publish y_i of every active scene -------------------------------------
        // A scene's y_i is the 16-byte (hi, lo) pair of words 2 s, 2 s + 1 of
        // record i, each word tagged, stored and read with single 16-byte
        // accesses at agent coherence (sc1, the cache policy of relaxed
        // agent-scope atomics); a pair seen half old and half new fails the
        // tag check and is read again.
        const int xb = static_cast<int>((tag & 1u) * static_cast<uint32_t>(K0) * kXW);   // word offset
        for (int e = tid; e < a.ipw * SMAX; e += NT) {
            const int t = e / SMAX, s = e - t * SMAX;
            double ys = s_ypart[0][t][s];
            s_ypart[0][t][s] = 0.0;
#pragma unroll
            for (int w = 1; w < kNW; ++w) {   // waves in order
                ys += s_ypart[w][t][s];
                s_ypart[w][t][s] = 0.0;
            }
            const int i = g + a.G * t;
            if (s < S && i < s_K[s]) {
                const unsigned long long bits = __double_as_longlong(ys);
                const u32x4 rec = {static_cast<uint32_t>(bits >> 32), tag, static_cast<uint32_t>(bits), tag};
                __builtin_amdgcn_raw_buffer_store_b128(rec, xrec, (xb + i * kXW + 2 * s) * 8, 0, kSc1);
            }
        }
        // ---- gather: whole records per wave load ------------------------------------
        // Lane l of wave wv reads the pair of scene l & 15 in record
        // 4 (wv + kNW r) + (l >> 4), r = 0, 1, ...: one wave-wide load covers four
        // adjacent records, 1 KiB contiguous.  Each pair lands in s_y[s][k].
        // Passes follow each other without a sleep.
        {
            const int sw = lane & 15, q = lane >> 4;
            const int Ksw = sw < S ? s_K[sw] : 0;
            uint32_t pend = 0;
#pragma unroll
            for (int r = 0; r < kPollQ; ++r)
                if (4 * (wv + kNW * r) + q < Ksw) pend |= 1u << r;
            // record 4 (wv + kNW r) + q, pair sw: a wave-uniform offset per r plus
            // the same lane offset for every r
            const int loff = xb + (4 * wv + q) * kXW + 2 * sw;
            double* ydst = s_y + sw * kSWS + 4 * wv + q;
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            bool timed_out = false;
            u32x4 vlast = {0u, 0u, 0u, 0u};
            while (pend) {
                u32x4 v[kPollQ];
#pragma unroll
                for (int r = 0; r < kPollQ; ++r) {
                    v[r] = u32x4{0u, 0u, 0u, 0u};
                    if ((pend >> r) & 1u)
                        v[r] = __builtin_amdgcn_raw_buffer_load_b128(xrec, (loff + r * (4 * kNW * kXW)) * 8, 0, kSc1);
                }
#pragma unroll
                for (int r = 0; r < kPollQ; ++r) {
                    if (((pend >> r) & 1u) && v[r].y == tag && v[r].w == tag) {
                        ydst[4 * kNW * r] =
                            __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(v[r].x) << 32) | v[r].z));
                        pend &= ~(1u << r);
                    }
                }
                if (pend && __builtin_amdgcn_s_memrealtime() - t0 > a.wait_ticks) {
                    timed_out = true;
#pragma unroll
                    for (int r = 0; r < kPollQ; ++r)
                        if (r == __builtin_ctz(pend)) vlast = v[r];
                    break;
                }
            }
            if (timed_out && atomicCAS(a.diag, 0ull, 1ull) == 0ull) {
                // the first wait of the launch to give up records what it waited for
                const int r = __builtin_ctz(pend);
                const bool hi_ok = vlast.y == tag;
                a.diag[1] = static_cast<unsigned long long>(a.instance) << 32 | static_cast<uint32_t>(g);
                a.diag[2] = static_cast<unsigned long long>(step) << 32 | static_cast<uint32_t>(a.nsteps);
                a.diag[3] = static_cast<unsigned long long>(4 * (wv + kNW * r) + q) << 32 | static_cast<uint32_t>(sw);
                a.diag[4] = tag;
                a.diag[5] = hi_ok ? (static_cast<unsigned long long>(vlast.w) << 32 | vlast.z)
                                  : (static_cast<unsigned long long>(vlast.y) << 32 | vlast.x);
                a.diag[6] = hi_ok ? 1ull : 0ull;
                a.diag[7] = static_cast<unsigned long long>(S) << 32 | static_cast<uint32_t>(a.G);
            }
            if (__any(timed_out) && lane == 0) s_flag = 2;
        }
        __syncthreads();
        // ---- update, one wave per scene (psfsolver.cpp:34-59) ----------------------------
        if (!s_flag) {
#pragma clang fp contract(off)
#pragma unroll
            for (int si = 0; si < SPW; ++si) {
                const int s = wv + kNW * si;
                if (!(s < S && step < s_n[s] && !s_dead[s])) continue;   // frozen: run its steps, or dead
                const int Ks = s_K[s];
                const GroupScene sc = a.scenes[s];
                double wn[JK];
                double s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int j = 0; j < JK; ++j) {
                    const int k = lane + 64 * j;
                    wn[j] = 0.0;
                    if (k < Ks) {
                        const double w = s_w[s * kSWS + k];
                        wn[j] = w + sc.dt * s_y[s * kSWS + k];
                        s1 += w * w;
                        s2 += wn[j] * wn[j];
                    }
                }
                const double e1 = wave_total(s1);
                const double e2 = wave_total(s2);
                if (!(e2 > 0.0) || !__builtin_isfinite(e2) || !__builtin_isfinite(e1)) {
                    if (lane == 0) s_dead[s] = 1;
                    continue;
                }
                const double fac = sqrt(e1 / e2);
#pragma unroll
                for (int j = 0; j < JK; ++j) {
                    const int k = lane + 64 * j;
                    if (k < Ks) {
                        double w = wn[j] * fac;
                        w = w * sc.visc[k];
                        if (sc.gravity_active) w = w + sc.dt * a.gravity[k];
                        s_w[s * kSWS + k] = w;
                        split_one(s, k, w);
                        if (sc.hist && g == 0) sc.hist[static_cast<size_t>(k) * sc.hist_ld + step] = w;
                    }
                }
            }
        }
        __syncthreads();
        if (s_flag) break;
    }
    __syncthreads();
    // every workgroup that gave up on the exchange reports it; workgroup 0
    // writes w and the energy status
    if (s_flag && tid < S) *a.scenes[tid].status = 2;
    if (g == 0) {
        for (int e = tid; e < S * kGroupMaxK; e += NT) {
            const int s = e / kGroupMaxK, k = e - s * kGroupMaxK;
            if (k < s_K[s]) a.scenes[s].omega[k] = s_w[s * kSWS + k];
        }
        if (tid < S && !s_flag && s_dead[tid]) *a.scenes[tid].status = 1;
    }
}

// ---------------------------------------------------------------------------
namespace {

// modes per workgroup: the most whose blocks fit the register budget
int group_ipw(int K0)
{
    const int nb = (K0 + kTile - 1) / kTile;
    const int nBi = nb * (nb + 1) / 2;
    const int ipw = (kNW * kGRB) / nBi;
    return std::max(1, std::min(ipw, kGMaxIpw));
}

template <int RB>
const void* kernel_for_ncs(int ncs)
{
    return ncs <= 1 ? reinterpret_cast<const void*>(&k_group_step<RB, 1>)
                    : reinterpret_cast<const void*>(&k_group_step<RB, 2>);
}

const void* pick_kernel(int per, int ncs)
{
    return per <= 12 ? kernel_for_ncs<12>(ncs)
           : per <= 20 ? kernel_for_ncs<20>(ncs)
           : per <= 30 ? kernel_for_ncs<30>(ncs)
                       : kernel_for_ncs<kGRB>(ncs);
}

}  // namespace

bool group_supported(int K0)
{
    if (K0 < 1 || K0 > kGroupMaxK) return false;
    int G, ipw, bpw;
    std::vector<uint32_t> meta;
    persistent_layout_meta(PSF_C_BF16X2, K0, group_ipw(K0), G, ipw, bpw, meta);
    return bpw <= kGMeta && ipw <= kGMaxIpw;
}

int group_basis_create(const double* Cf, int K0, const double* gravity, hipStream_t stream, GroupBasis** out,
                       std::string* err)
{
    *out = nullptr;
    if (!group_supported(K0)) {
        if (err) *err = "basis K0 " + std::to_string(K0) + " has no shared-basis layout (K0 <= 256)";
        return PSF_ERR_UNSUPPORTED;
    }
    GroupBasis* b = new (std::nothrow) GroupBasis();
    if (!b) return PSF_ERR_NOMEM;
    std::vector<uint32_t> meta;
    persistent_layout_meta(PSF_C_BF16X2, K0, group_ipw(K0), b->G, b->ipw, b->bpw, meta);
    b->K0 = K0;
    b->gravity = gravity;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->blocks), static_cast<size_t>(b->G) * b->bpw * 64 * 16);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->meta), sizeof(uint32_t) * b->bpw);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->inv_scale), sizeof(double) * b->G * b->ipw);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->meta, meta.data(), sizeof(uint32_t) * meta.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = launch_pack_c_layout(Cf, K0, K0, b->G, b->ipw, b->bpw, 1, b->meta, b->blocks, b->inv_scale, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        if (err) *err = std::string("shared-basis blocks: ") + hipGetErrorString(e);
        group_basis_destroy(b);
        return e == hipErrorOutOfMemory ? PSF_ERR_NOMEM : PSF_ERR_HIP;
    }
    *out = b;
    return PSF_OK;
}

void group_basis_destroy(GroupBasis* b)
{
    if (!b) return;
    if (b->blocks) (void)hipFree(b->blocks);
    if (b->meta) (void)hipFree(b->meta);
    if (b->inv_scale) (void)hipFree(b->inv_scale);
    delete b;
}

int group_basis_workgroups(const GroupBasis* b) { return b->G; }
size_t group_basis_bytes(const GroupBasis* b) { return static_cast<size_t>(b->G) * b->bpw * 64 * 16; }

GroupBatch* group_batch_create(int device, int max_instances)
{
    GroupBatch* b = new (std::nothrow) GroupBatch();
    if (!b) return nullptr;
    int cus = 0, per_cu = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_group_step<kGRB, 2>),
                                                         64 * kNW, 0);
    b->max_instances = max_instances;
    b->capacity = per_cu * cus;
    b->max_wgs = std::max(b->capacity, 1);
    const size_t slot = static_cast<size_t>(2) * kGroupMaxK * kXW;
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_args), sizeof(GroupArgs) * max_instances);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void**>(&b->d_scenes), sizeof(GroupScene) * max_instances * kGroupMaxScenes);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_map), sizeof(uint32_t) * b->max_wgs);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&b->h_args), sizeof(GroupArgs) * max_instances, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&b->h_scenes), sizeof(GroupScene) * max_instances * kGroupMaxScenes,
                          hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_map), sizeof(uint32_t) * b->max_wgs, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->xbuf), slot * max_instances * 8);
    if (e == hipSuccess) e = hipMemset(b->xbuf, 0, slot * max_instances * 8);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->diag), kDiagWords * 8);
    if (e == hipSuccess) e = hipMemset(b->diag, 0, kDiagWords * 8);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess || b->capacity < 1) {
        group_batch_destroy(b);
        return nullptr;
    }
    b->tags.assign(max_instances, 0u);
    return b;
}

void group_batch_destroy(GroupBatch* b)
{
    if (!b) return;
    if (b->d_args) (void)hipFree(b->d_args);
    if (b->d_scenes) (void)hipFree(b->d_scenes);
    if (b->d_map) (void)hipFree(b->d_map);
    if (b->h_args) (void)hipHostFree(b->h_args);
    if (b->h_scenes) (void)hipHostFree(b->h_scenes);
    if (b->h_map) (void)hipHostFree(b->h_map);
    if (b->xbuf) (void)hipFree(b->xbuf);
    if (b->diag) (void)hipFree(b->diag);
    delete b;
}

int group_batch_capacity(const GroupBatch* b) { return b->capacity; }

std::string group_batch_diag(GroupBatch* b)
{
    unsigned long long d[kDiagWords] = {};
    if (hipMemcpy(d, b->diag, sizeof(d), hipMemcpyDeviceToHost) != hipSuccess || d[0] == 0) return std::string();
    (void)hipMemset(b->diag, 0, sizeof(d));
    char buf[320];
    snprintf(buf, sizeof(buf),
             "instance %llu (%llu scenes, %llu workgroups), workgroup %llu, step %llu of %llu: mode %llu of scene "
             "%llu, tag %llu expected, %s word 0x%016llx",
             d[1] >> 32, d[7] >> 32, d[7] & 0xFFFFFFFFull, d[1] & 0xFFFFFFFFull, d[2] >> 32, d[2] & 0xFFFFFFFFull,
             d[3] >> 32, d[3] & 0xFFFFFFFFull, d[4], d[6] ? "lo" : "hi", d[5]);
    return std::string(buf);
}
int group_batch_max_instances(const GroupBatch* b) { return b->max_instances; }

hipError_t launch_group_steps(GroupBatch* b, const GroupInstance* inst, int n_inst, uint64_t wait_ticks,
                              hipStream_t stream)
{
    if (n_inst <= 0) return hipSuccess;
    if (n_inst > b->max_instances) return hipErrorInvalidValue;
    const size_t slot = static_cast<size_t>(2) * kGroupMaxK * kXW;
    int wgs = 0, per_max = 0, ncs = 1;
    std::vector<int> nsteps(n_inst, 0), gl(n_inst, 0);
    for (int j = 0; j < n_inst; ++j) {
        const GroupInstance& I = inst[j];
        if (!I.basis || I.n < 1 || I.n > kGroupMaxScenes) return hipErrorInvalidValue;
        int kmax = 0;
        for (int s = 0; s < I.n; ++s) {
            if (I.scenes[s].K < 1 || I.scenes[s].K > I.basis->K0 || I.scenes[s].nsteps < 0) return hipErrorInvalidValue;
            nsteps[j] = std::max(nsteps[j], I.scenes[s].nsteps);
            kmax = std::max(kmax, I.scenes[s].K);
        }
        if (nsteps[j] == 0) continue;
        gl[j] = std::min(I.basis->G, kmax);   // the workgroups owning a mode below the largest K
        wgs += gl[j];
        per_max = std::max(per_max, (I.basis->bpw + kNW - 1) / kNW);
        ncs = std::max(ncs, (I.n + 7) / 8);
    }
    if (wgs == 0) return hipSuccess;
    if (wgs > b->capacity) return hipErrorCooperativeLaunchTooLarge;
    if (per_max > kGRB) return hipErrorInvalidValue;
    int w = 0, na = 0;
    for (int j = 0; j < n_inst; ++j) {
        if (nsteps[j] == 0) continue;
        const GroupInstance& I = inst[j];
        GroupScene* sc = b->h_scenes + static_cast<size_t>(na) * kGroupMaxScenes;
        std::memcpy(sc, I.scenes, sizeof(GroupScene) * I.n);
        GroupArgs& A = b->h_args[na];
        A.blocks = I.basis->blocks;
        A.meta = I.basis->meta;
        A.gravity = I.basis->gravity;
        A.xbuf = b->xbuf + slot * na;
        A.scenes = b->d_scenes + static_cast<size_t>(na) * kGroupMaxScenes;
        A.diag = b->diag;
        A.instance = na;
        A.S = I.n;
        A.K0 = I.basis->K0;
        A.G = I.basis->G;
        A.ipw = I.basis->ipw;
        A.bpw = I.basis->bpw;
        A.nsteps = nsteps[j];
        A.tag0 = b->tags[na];
        A.wait_ticks = wait_ticks ? wait_ticks : kGWait;
        for (int g = 0; g < gl[j]; ++g) b->h_map[w++] = static_cast<uint32_t>(na) << 16 | static_cast<uint32_t>(g);
        ++na;
    }
    hipError_t e = hipMemcpyAsync(b->d_args, b->h_args, sizeof(GroupArgs) * na, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_scenes, b->h_scenes, sizeof(GroupScene) * na * kGroupMaxScenes, hipMemcpyHostToDevice,
                           stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_map, b->h_map, sizeof(uint32_t) * wgs, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    const GroupArgs* targ = b->d_args;
    const uint32_t* tmap = b->d_map;
    void* args[] = {&targ, &tmap};
    e = hipLaunchKernel(pick_kernel(per_max, ncs), dim3(wgs), dim3(64 * kNW), args, 0, stream);
    if (e != hipSuccess) return e;
    na = 0;
    for (int j = 0; j < n_inst; ++j)
        if (nsteps[j] > 0) b->tags[na++] += static_cast<uint32_t>(nsteps[j]);
    return hipSuccess;
}

}  // namespace psf
