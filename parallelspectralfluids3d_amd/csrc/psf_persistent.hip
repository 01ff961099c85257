// Persistent reduced-step path on MFMA (BF16, FP16, BF16 hi+lo).
//
// One launch runs n steps of psfsolver.cpp:34-59.
//
// Storage.  Only the symmetric part of C_i enters y_i = w^T C_i w, so C_i is
// replaced by S_i = (C_i + C_i^T)/2 and stored as its 16x16 tiles T_pq with
// p <= q (T_pp = S_pp, T_pq = 2 S_pq): y_i = sum_{p<=q} w_p^T T_pq w_q.
// An MFMA block is the A operand of one v_mfma_f32_16x16x32: 16 rows (the
// row tile p) by 32 reduction entries, either two column tiles (q1, q2) of a
// single-precision tensor or the hi and lo parts of one tile.  Blocks are
// stored pre-swizzled in the per-lane register layout (16 bytes per lane, a
// coalesced 1 KiB load per wave) and are loaded into VGPRs once per launch.
//
// w enters as a pair of B columns: one holds w_hi, the next w_lo (w - w_hi,
// scaled for FP16), so one MFMA gives T (w_hi + w_lo) with FP32 accumulation.
// A wave's resident blocks cover a few row groups (runs of blocks with the
// same mode and row tile); row group s of the wave takes column pair
// (2s, 2s + 1) and every other column of its B operand is zero, so the whole
// chain accumulates without a break and the row groups are separated only
// once, after the last MFMA.  The row contraction with w_p and all
// reductions over rows, blocks, waves and workgroups are FP64.
//
// Exchange.  Workgroup g owns modes i = g, g+G, ...  After the contraction it
// publishes y_i as a 16-byte record of two 64-bit words (step tag << 32 |
// high or low half of the FP64 bits), stored and polled with single 16-byte
// buffer accesses at agent coherence (sc1), and one wave of every workgroup
// polls all K records until both tags match.  Each word carries its own tag
// in the same aligned 8-byte unit, so a record seen half old and half new
// fails the check and is read again.  Each mode's record sits on its own
// 128-byte line: records that share a line written from several CUs cost
// more than reading one line per mode.  No release/acquire fences (L2
// write-back/invalidate) are needed, and the y values arrive exactly.
// Buffers alternate by tag parity: a workgroup can only overwrite a buffer
// two steps later, after every workgroup has consumed it.  Every workgroup
// then applies the same FP64 update (energy renormalisation, viscosity,
// forcing) to its own copy of w, in a fixed order, so all copies stay
// bitwise identical; the energy of w before the step is summed by a second
// wave while the first one polls.  Waits are bounded (2 s of device time);
// on expiry every workgroup leaves the loop and the status word reports it.
//
// Residency.  The exchange needs all G workgroups resident at once.  The
// launch is an ordinary one on the scene's stream: kernels on one stream run
// one after another, so the GPU holds no other work of the scene when the
// step starts, and scene set-up checks that G workgroups fit on the device
// at the kernel's occupancy.  Should a workgroup still not be scheduled, the
// bounded waits turn that into the timed-out status, not a hang.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <hip/hip_ext.h>

#include "psf_persistent.hpp"
#include "psf_wave.hpp"

namespace psf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 16;
constexpr int kRB = 36;          // largest register budget, in MFMA blocks per wave (K = 256 hi+lo needs 34)
constexpr int kMaxK = 1024;      // LDS capacity for w, y
constexpr int kMaxMeta = 4096;   // blocks per workgroup
constexpr int kMaxIpw = 8;       // modes per workgroup
constexpr int kMaxCUs = 256;
constexpr uint64_t kWaitTicks = 200000000ull;   // s_memrealtime runs at 100 MHz: 2 s
constexpr int kXStride = 16;     // 64-bit words per exchange record: one 128-byte line per mode
constexpr int kBChunk = 6;       // B-operand blocks loaded ahead of the MFMA chain
// waves per workgroup of a single scene's launch (one workgroup per CU): two
// per SIMD, so one wave's MFMA, LDS and FP64 latencies overlap the other's
#ifndef PSF_SCENE_WAVES
#define PSF_SCENE_WAVES 8
#endif
constexpr int kSceneWaves = PSF_SCENE_WAVES;

struct PersistentState {
    int G = 0, ipw = 0, Kp = 0, nb = 0, nBi = 0, bpw = 0;
    int elem = 0;                // 0 bf16, 1 fp16
    float col1 = 1.0f;           // scale of the w_lo column
    float wscale = 1.0f;         // w is multiplied by this before FP16 rounding
    uint4* blocks = nullptr;     // [G][bpw][64]
    uint32_t* meta = nullptr;    // [bpw]: t << 24 | p << 16 | q1 << 8 | q2
    double* inv_scale = nullptr; // [G * ipw]
    unsigned long long* xbuf = nullptr;   // [2][K][kXStride], words 0 and 1 used
    uint32_t tag = 0;
    uint64_t wait_ticks = kWaitTicks;   // bound of each exchange wait (s_memrealtime ticks)
    const void* fn = nullptr;    // the k_persistent instance for this layout
};

struct PArgs {
    const uint4* blocks;
    const uint32_t* meta;
    const double* inv_scale;
    unsigned long long* xbuf;
    double* omega;
    const double* visc;
    const double* gravity;
    int* status;
    double dt;
    int gravity_active;
    int K, Kp, G, ipw, bpw, nsteps;
    uint32_t tag0;
    float col1, wscale;
    double* hist;                // optional: w after step s at hist[k * hist_ld + s]
    int hist_ld;
    uint64_t wait_ticks;         // bound of each exchange wait
};

static inline PersistentState* st(psf_scene* s) { return reinterpret_cast<PersistentState*>(s->persistent); }

// Row tiles of a mode in storage order: 0, nb-1, 1, nb-2, ...  Row p holds
// nb - p tile blocks (about half as many for two-tile blocks), so a long row
// and a short row alternate; the contiguous block ranges of the waves then
// hold nearly equal numbers of row groups (two each at K = 256 with eight
// waves), and each group takes a column pair of the wave's accumulator.
static inline int row_at(int idx, int nb) { return (idx & 1) ? nb - 1 - (idx >> 1) : (idx >> 1); }

int blocks_per_mode(int prec, int nb)
{
    if (prec == PSF_C_BF16X2) return nb * (nb + 1) / 2;
    int n = 0;
    for (int p = 0; p < nb; ++p) n += (nb - p + 1) / 2;
    return n;
}

// Workgroup g owns modes g, g + G, ...  By default one mode per workgroup up
// to kMaxCUs workgroups; ipw_req > 0 asks for that many modes per workgroup
// (fewer workgroups per scene, so more scenes fit in one launch).
static void layout(int prec, int K, int ipw_req, int& G, int& ipw, int& nb, int& nBi, int& bpw)
{
    G = ipw_req > 0 ? (K + ipw_req - 1) / ipw_req : (K < kMaxCUs ? K : kMaxCUs);
    ipw = (K + G - 1) / G;
    nb = (K + kTile - 1) / kTile;
    nBi = blocks_per_mode(prec, nb);
    bpw = ipw * nBi;
}

bool persistent_supported(int prec, int K, int ipw_req)
{
    if (prec != PSF_C_BF16 && prec != PSF_C_FP16 && prec != PSF_C_BF16X2) return false;
    if (K < 1 || K > kMaxK || ipw_req < 0 || ipw_req > kMaxIpw) return false;
    int G, ipw, nb, nBi, bpw;
    layout(prec, K, ipw_req, G, ipw, nb, nBi, bpw);
    return ipw <= kMaxIpw && bpw <= kMaxMeta && G <= 65535;
}

size_t persistent_c_bytes(int prec, int K, int ipw_req)
{
    if (!persistent_supported(prec, K, ipw_req)) return 0;
    int G, ipw, nb, nBi, bpw;
    layout(prec, K, ipw_req, G, ipw, nb, nBi, bpw);
    return size_t(G) * bpw * 64 * 16;
}

int persistent_auto_ipw(int prec, int K)
{
    const int nb = (K + kTile - 1) / kTile;
    const int nBi = blocks_per_mode(prec, nb);
    int ipw = (4 * kRB) / nBi;
    return ipw < 1 ? 1 : (ipw > kMaxIpw ? kMaxIpw : ipw);
}

int persistent_workgroups(int prec, int K, int ipw_req)
{
    if (!persistent_supported(prec, K, ipw_req)) return 0;
    int G, ipw, nb, nBi, bpw;
    layout(prec, K, ipw_req, G, ipw, nb, nBi, bpw);
    return G;
}

template <int ELEM>
__device__ __forceinline__ f32x4 mfma16(const uint4& a, const uint4& b, const f32x4& c)
{
    if constexpr (ELEM == 0)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// The value of the lane 2 S columns to the right in the same row of 16
// (DPP row_shl); S is a compile-time constant after unrolling.
__device__ __forceinline__ double from_pair(double v, int S)
{
    switch (S) {
        case 1: return dpp_move<0x102, 0xF>(v);
        case 2: return dpp_move<0x104, 0xF>(v);
        case 3: return dpp_move<0x106, 0xF>(v);
        case 4: return dpp_move<0x108, 0xF>(v);
        case 5: return dpp_move<0x10A, 0xF>(v);
        case 6: return dpp_move<0x10C, 0xF>(v);
        default: return dpp_move<0x10E, 0xF>(v);
    }
}

// The body of one workgroup (NW waves): local workgroup g of the scene
// described by a.
template <int ELEM, int RB, int NW>
__device__ __forceinline__ void persistent_body(const PArgs& a, const int g)
{
    constexpr int NT = 64 * NW;
    __shared__ double s_w[kMaxK];
    __shared__ double s_y[kMaxK];
    __shared__ __align__(16) uint16_t s_wb[3 * kMaxK];   // w_hi at [0, Kp), w_lo at [Kp, 2 Kp), zeros after
    __shared__ uint32_t s_meta[kMaxMeta];
    __shared__ double s_visc[kMaxK];
    __shared__ double s_grav[kMaxK];
    __shared__ double s_red[NW + 1];   // per-wave sums of the new energy, then the old energy
    __shared__ double s_ypart[NW][kMaxIpw];
    __shared__ int s_flag;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int K = a.K, Kp = a.Kp;
    const int col = lane & 15, kg = lane >> 4;

    for (int b = tid; b < a.bpw; b += NT) s_meta[b] = a.meta[b];
    for (int k = tid; k < Kp; k += NT) {
        s_w[k] = k < K ? a.omega[k] : 0.0;
        s_visc[k] = k < K ? a.visc[k] : 1.0;
        s_grav[k] = k < K ? a.gravity[k] : 0.0;
    }
    for (int k = tid; k < Kp; k += NT) s_wb[2 * Kp + k] = 0;
    if (tid == 0) s_flag = 0;

    // this wave's block range; the first RB blocks stay in registers, and
    // register slots past the wave's last block hold zero blocks, so the
    // resident MFMA chain runs without guards
    const int per = (a.bpw + NW - 1) / NW;
    const int b0 = wv * per;
    const int nblk = max(0, min(a.bpw - b0, per));
    const int nres = min(nblk, RB);
    const uint4* gblk = a.blocks + (static_cast<size_t>(g) * a.bpw + b0) * 64 + lane;
    uint4 areg[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) {
        areg[j] = uint4{0u, 0u, 0u, 0u};
        if (j < nres) areg[j] = gblk[static_cast<size_t>(j) * 64];
    }
    __syncthreads();
    // metadata of the resident blocks, one word per lane, read back with readlane;
    // fmask bit j: block j ends a row group, tmask bit j: it ends a mode
    const uint32_t mreg = lane < nres ? s_meta[b0 + lane] : 0u;
    uint64_t fmask = 0, tmask = 0;
    for (int j = 0; j < nres; ++j) {
        const int kj = __builtin_amdgcn_readlane(static_cast<int>(mreg), j) >> 16;
        const int kn = j + 1 < nres ? (__builtin_amdgcn_readlane(static_cast<int>(mreg), j + 1) >> 16) : -1;
        if (kn != kj) fmask |= 1ull << j;
        if (kn < 0 || (kn >> 8) != (kj >> 8)) tmask |= 1ull << j;
    }
    // Column pairs: with at most 8 row groups and every block resident, row
    // group s accumulates into columns (2s, 2s + 1) over the whole chain.
    // Otherwise the chain flushes columns 0 and 1 at the end of each group;
    // so do the chains of the largest register budget, where the pair form's
    // addresses would spill.
    const int nseg = __builtin_popcountll(fmask);
    const bool pairs = RB <= 18 && nseg <= 8 && nblk <= RB;
    const int lane_seg = col >> 1;
    const int bsel_on = ((col & 1) ? Kp : 0) + (kg & 1) * 8;   // this lane's half tile of w_hi or w_lo
    const int bsel_off = 2 * Kp + (kg & 1) * 8;                // the zeroed region
    // per lane: row tile of its column pair's group; per wave: the mode of
    // each group (4 bits each) and which groups end a mode
    int lane_p = 0;
    uint32_t seg_t = 0, seg_end = 0;
    {
        int k = 0, jend = 0;
        for (int j = 0; j < nres; ++j)
            if ((fmask >> j) & 1ull) {
                if (k == lane_seg) jend = j;
                const uint32_t mj = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), j));
                if (k < 8) {
                    seg_t |= ((mj >> 24) & 15u) << (4 * k);
                    if ((tmask >> j) & 1ull) seg_end |= 1u << k;
                }
                ++k;
            }
        lane_p = static_cast<int>((static_cast<uint32_t>(__shfl(static_cast<int>(mreg), jend)) >> 16) & 255u);
    }
    // B operand of block j (metadata m): in pair form the lanes of the block's
    // column pair read w_hi / w_lo and the others the zeroed region, so every
    // lane issues the same unpredicated load
    auto b_pair = [&](int j) {
        const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), j));
        const int sj = __builtin_popcountll(fmask & ((1ull << j) - 1ull));
        const int q = (kg < 2) ? static_cast<int>((m >> 8) & 255u) : static_cast<int>(m & 255u);
        return *reinterpret_cast<const uint4*>(s_wb + (lane_seg == sj ? bsel_on : bsel_off) + q * kTile);
    };
    const int bsel = col < 2 ? bsel_on : bsel_off;
    auto b_operand = [&](uint32_t m) {
        const int q = (kg < 2) ? static_cast<int>((m >> 8) & 255u) : static_cast<int>(m & 255u);
        return *reinterpret_cast<const uint4*>(s_wb + bsel + q * kTile);
    };
    // FP64 contraction of a 16 x 16 accumulator's column with w over row tile p
    auto col_contract = [&](const f32x4& acc, int p) {
        const double* wr = s_w + p * kTile + 4 * kg;
        double c = static_cast<double>(acc[0]) * wr[0] + static_cast<double>(acc[1]) * wr[1] +
                   static_cast<double>(acc[2]) * wr[2] + static_cast<double>(acc[3]) * wr[3];
        if (col & 1) c *= static_cast<double>(a.col1);
        return c;
    };
    auto row_contract = [&](const f32x4& acc, int p) { return col < 2 ? col_contract(acc, p) : 0.0; };

    // hi/lo split of one coefficient into the B-operand table
    auto split_one = [&](int k, double w) {
        uint16_t hi, lo;
        if constexpr (ELEM == 0) {
            hi = f32_to_bf16_rne(static_cast<float>(w));
            lo = f32_to_bf16_rne(static_cast<float>(w - static_cast<double>(bf16_to_f32(hi))));
        } else {
            const double ws = w * static_cast<double>(a.wscale);
            const _Float16 h = static_cast<_Float16>(static_cast<float>(ws));
            const _Float16 l = static_cast<_Float16>(static_cast<float>((ws - static_cast<double>(h)) * 2048.0));
            hi = __builtin_bit_cast(uint16_t, h);
            lo = __builtin_bit_cast(uint16_t, l);
        }
        s_wb[k] = hi;
        s_wb[Kp + k] = lo;
    };
    for (int k = tid; k < Kp; k += NT) split_one(k, s_w[k]);
    if (tid < NW * kMaxIpw) s_ypart[tid / kMaxIpw][tid % kMaxIpw] = 0.0;
    __syncthreads();
    // the exchange records as a buffer: 16-byte accesses with agent coherence
    // (sc1, the cache policy of relaxed agent-scope atomics); xbuf holds
    // 2 K records of kXStride words
    constexpr int kSc1 = 16;
    const __amdgpu_buffer_rsrc_t xrec = __builtin_amdgcn_make_buffer_rsrc(
        a.xbuf, 0, static_cast<int>(sizeof(unsigned long long) * 2 * kXStride * K), 0x00020000);

    for (int s = 0; s < a.nsteps; ++s) {
        const uint32_t tag = a.tag0 + static_cast<uint32_t>(s) + 1u;
        // ---- contraction of this wave's blocks -------------------------------
        // B operands come from LDS in chunks of kBChunk blocks, the next chunk
        // loading while the current one feeds the MFMAs
        const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
        if (RB <= 18 && pairs) {
            uint4 bq[2][kBChunk];
#pragma unroll
            for (int jj = 0; jj < kBChunk && jj < RB; ++jj) bq[0][jj] = b_pair(jj);
            f32x4 acc0 = zero4, acc1 = zero4;
#pragma unroll
            for (int c = 0; c < RB; c += kBChunk) {
                const int cb = (c / kBChunk) & 1;
#pragma unroll
                for (int jj = 0; jj < kBChunk; ++jj)
                    if (c + kBChunk + jj < RB) bq[cb ^ 1][jj] = b_pair(c + kBChunk + jj);
#pragma unroll
                for (int jj = 0; jj < kBChunk; ++jj) {
                    const int j = c + jj;
                    if (j >= RB) break;
                    if (j & 1) acc1 = mfma16<ELEM>(areg[j], bq[cb][jj], acc1);
                    else acc0 = mfma16<ELEM>(areg[j], bq[cb][jj], acc0);
                }
            }
            // each lane contracts its column with its group's row tile; the
            // groups then come together in order in columns 0 and 1, one wave
            // sum per mode
            const double cv = lane_seg < nseg ? col_contract(acc0 + acc1, lane_p) : 0.0;
            double part = cv;
            auto flush_mode = [&](int sg) {
                const double v = wave_total(col < 2 ? part : 0.0);
                if (lane == 0) s_ypart[wv][(seg_t >> (4 * sg)) & 15u] += v;
            };
#pragma unroll
            for (int sg = 1; sg < 8; ++sg) {
                if (sg >= nseg) break;
                const double cs = from_pair(cv, sg);
                if ((seg_end >> (sg - 1)) & 1u) {
                    flush_mode(sg - 1);
                    part = cs;
                } else {
                    part += cs;
                }
            }
            if (nseg > 0) flush_mode(nseg - 1);
        } else {
            uint4 bq[2][kBChunk];
#pragma unroll
            for (int jj = 0; jj < kBChunk && jj < RB; ++jj)
                bq[0][jj] = b_operand(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), jj)));
            f32x4 acc0 = zero4, acc1 = zero4;
            double part = 0.0;
#pragma unroll
            for (int c = 0; c < RB; c += kBChunk) {
                const int cb = (c / kBChunk) & 1;
#pragma unroll
                for (int jj = 0; jj < kBChunk; ++jj)
                    if (c + kBChunk + jj < RB)
                        bq[cb ^ 1][jj] = b_operand(
                            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), c + kBChunk + jj)));
#pragma unroll
                for (int jj = 0; jj < kBChunk; ++jj) {
                    const int j = c + jj;
                    if (j >= RB) break;
                    if (j & 1) acc1 = mfma16<ELEM>(areg[j], bq[cb][jj], acc1);
                    else acc0 = mfma16<ELEM>(areg[j], bq[cb][jj], acc0);
                    if ((fmask >> j) & 1ull) {
                        const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mreg), j));
                        part += row_contract(acc0 + acc1, static_cast<int>((m >> 16) & 255u));
                        acc0 = zero4;
                        acc1 = zero4;
                        if ((tmask >> j) & 1ull) {
                            const double v = wave_total(part);
                            if (lane == 0) s_ypart[wv][m >> 24] += v;
                            part = 0.0;
                        }
                    }
                }
            }
            // blocks beyond the register budget stream from HBM every step
            int cur = -1;
            for (int j = RB; j < nblk; ++j) {
                const uint32_t m = s_meta[b0 + j];
                const int key = static_cast<int>(m >> 16);
                if (key != cur) {
                    if (cur >= 0) {
                        part += row_contract(acc0, cur & 255);
                        if ((key >> 8) != (cur >> 8)) {
                            const double v = wave_total(part);
                            if (lane == 0) s_ypart[wv][cur >> 8] += v;
                            part = 0.0;
                        }
                    }
                    acc0 = zero4;
                    cur = key;
                }
                acc0 = mfma16<ELEM>(gblk[static_cast<size_t>(j) * 64], b_operand(m), acc0);
            }
            if (cur >= 0) {
                part += row_contract(acc0, cur & 255);
                const double v = wave_total(part);
                if (lane == 0) s_ypart[wv][cur >> 8] += v;
            }
        }
        __syncthreads();
        // ---- publish y_i of this workgroup ---------------------------------------
        const int xb = static_cast<int>((tag & 1u) * static_cast<uint32_t>(K) * kXStride);   // word offset
        if (tid < a.ipw) {
            const int i = g + a.G * tid;
            if (i < K) {
                double ys = s_ypart[0][tid];
#pragma unroll
                for (int w = 1; w < NW; ++w) ys += s_ypart[w][tid];   // waves in order
                const double y = ys * a.inv_scale[g * a.ipw + tid];
                const unsigned long long bits = __double_as_longlong(y);
                const u32x4 rec = {static_cast<uint32_t>(bits >> 32), tag, static_cast<uint32_t>(bits), tag};
                __builtin_amdgcn_raw_buffer_store_b128(rec, xrec, (xb + kXStride * i) * 8, 0, kSc1);
            }
        }
        // ---- energy of w before the step, while wave 0 polls ------------------------
        if (wv == NW - 1) {
#pragma clang fp contract(off)
            double s1 = 0.0;
            for (int k = lane; k < K; k += 64) s1 += s_w[k] * s_w[k];
            s1 = wave_total(s1);
            if (lane == 0) s_red[NW] = s1;
        }
        // ---- gather all y (one wave polls; bounded wait) ----------------------------
        // Lane l reads records l, l + 64, ..., four at a time: a poll pass
        // issues the loads of all still-missing records of the group before
        // checking any, so for K <= 256 a pass costs one round trip to the
        // other XCDs' writes, not one per record.  Passes follow each other
        // without a sleep (measured 2% more steps/s than sleeping between them).
        if (wv == 0) {
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            constexpr int kGroup = 4;
            const int nr = lane < K ? (K - 1 - lane) / 64 + 1 : 0;
            for (int r0 = 0; r0 < nr && !s_flag; r0 += kGroup) {
                uint32_t pend = (1u << min(kGroup, nr - r0)) - 1u;
                while (pend) {
                    u32x4 v[kGroup];
#pragma unroll
                    for (int r = 0; r < kGroup; ++r) {
                        v[r] = u32x4{0u, 0u, 0u, 0u};
                        if ((pend >> r) & 1u)
                            v[r] = __builtin_amdgcn_raw_buffer_load_b128(
                                xrec, (xb + kXStride * (lane + 64 * (r0 + r))) * 8, 0, kSc1);
                    }
#pragma unroll
                    for (int r = 0; r < kGroup; ++r) {
                        if (((pend >> r) & 1u) && v[r].y == tag && v[r].w == tag) {
                            s_y[lane + 64 * (r0 + r)] = __longlong_as_double(
                                static_cast<long long>((static_cast<unsigned long long>(v[r].x) << 32) | v[r].z));
                            pend &= ~(1u << r);
                        }
                    }
                    if (pend) {
                        if (__builtin_amdgcn_s_memrealtime() - t0 > a.wait_ticks) {
                            s_flag = 2;
                            break;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (s_flag) break;
        // ---- FP64 update, identical in every workgroup (psfsolver.cpp:34-59) -----------
        {
#pragma clang fp contract(off)
            double s2 = 0.0;
            for (int k = tid; k < K; k += NT) {
                const double wn = s_w[k] + a.dt * s_y[k];
                s_y[k] = wn;
                s2 += wn * wn;
            }
            s2 = wave_total(s2);
            if (lane == 0) s_red[wv] = s2;
            __syncthreads();
            const double e1 = s_red[NW];
            double e2 = s_red[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) e2 += s_red[w];   // waves in order
            if (!(e2 > 0.0) || !__builtin_isfinite(e2) || !__builtin_isfinite(e1)) {
                if (tid == 0) s_flag = 1;
                __syncthreads();
                break;
            }
            const double fac = sqrt(e1 / e2);
            for (int k = tid; k < K; k += NT) {
                double w = s_y[k] * fac;
                w = w * s_visc[k];
                if (a.gravity_active) w = w + a.dt * s_grav[k];
                s_w[k] = w;
                split_one(k, w);
                if (a.hist && g == 0) a.hist[static_cast<size_t>(k) * a.hist_ld + s] = w;
            }
            if (tid < NW * kMaxIpw) s_ypart[tid / kMaxIpw][tid % kMaxIpw] = 0.0;
        }
        __syncthreads();
    }
    __syncthreads();
    if (g == 0) {
        for (int k = tid; k < K; k += NT) a.omega[k] = s_w[k];
        if (tid == 0 && s_flag) *a.status = s_flag;
    }
}

// one scene: workgroup g = blockIdx.x, NW waves each
template <int ELEM, int RB, int NW>
__global__ __launch_bounds__(64 * NW, 1) void k_persistent(PArgs a)
{
    persistent_body<ELEM, RB, NW>(a, static_cast<int>(blockIdx.x));
}

// several independent scenes in one launch: wg_map[blockIdx.x] = scene << 16 | g.
// Each scene's workgroups exchange only among themselves (own xbuf, tags,
// step count), so a scene gives the same bits alone and in a batch.
template <int ELEM, int RB>
__global__ __launch_bounds__(256, 2) void k_persistent_multi(const PArgs* __restrict__ table,
                                                             const uint32_t* __restrict__ wg_map)
{
    const uint32_t m = wg_map[blockIdx.x];
    const PArgs a = table[m >> 16];
    persistent_body<ELEM, RB, 4>(a, static_cast<int>(m & 0xFFFFu));
}

// ---------------------------------------------------------------------------
// smallest register budget that holds a wave's blocks (kRB when they do not fit)
template <int ELEM>
static const void* pick_kernel_elem(int bpw)
{
    constexpr int NW = kSceneWaves;
    const int per = (bpw + NW - 1) / NW;
    return per <= 4    ? reinterpret_cast<const void*>(&k_persistent<ELEM, 4, NW>)
           : per <= 12 ? reinterpret_cast<const void*>(&k_persistent<ELEM, 12, NW>)
           : per <= 18 ? reinterpret_cast<const void*>(&k_persistent<ELEM, 18, NW>)
                       : reinterpret_cast<const void*>(&k_persistent<ELEM, kRB, NW>);
}

static const void* pick_kernel(int elem, int bpw)
{
    return elem == 0 ? pick_kernel_elem<0>(bpw) : pick_kernel_elem<1>(bpw);
}

hipError_t persistent_scene_init(psf_scene* s)
{
    const int prec = s->d.c_precision;
    if (!persistent_supported(prec, static_cast<int>(s->d.K), s->modes_per_wg)) return hipSuccess;
    PersistentState* p = new (std::nothrow) PersistentState();
    if (!p) return hipErrorOutOfMemory;
    s->persistent = p;
    const int K = static_cast<int>(s->d.K);
    layout(prec, K, s->modes_per_wg, p->G, p->ipw, p->nb, p->nBi, p->bpw);
    p->Kp = p->nb * kTile;
    p->elem = prec == PSF_C_FP16 ? 1 : 0;
    p->col1 = prec == PSF_C_FP16 ? 1.0f / 2048.0f : 1.0f;
    p->wscale = prec == PSF_C_FP16 ? 256.0f : 1.0f;
    p->blocks = static_cast<uint4*>(s->dev.C);
    p->fn = pick_kernel(p->elem, p->bpw);
    // every workgroup must be resident at once
    int per_cu = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p->fn, 64 * kSceneWaves, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->ctx->device);
    if (e != hipSuccess) return e;
    if (per_cu * cus < p->G) return hipErrorCooperativeLaunchTooLarge;
    e = scene_alloc(s, reinterpret_cast<void**>(&p->meta), sizeof(uint32_t) * p->bpw);
    if (e == hipSuccess) e = scene_alloc(s, reinterpret_cast<void**>(&p->inv_scale), sizeof(double) * p->G * p->ipw);
    if (e == hipSuccess) e = scene_alloc(s, reinterpret_cast<void**>(&p->xbuf), sizeof(unsigned long long) * 2 * kXStride * K);
    if (e == hipSuccess) e = hipMemset(p->xbuf, 0, sizeof(unsigned long long) * 2 * kXStride * K);
    if (e != hipSuccess) return e;
    // block metadata (identical for every workgroup)
    std::vector<uint32_t> meta;
    int G = 0, ipw = 0, bpw = 0;
    persistent_layout_meta(prec, K, s->modes_per_wg, G, ipw, bpw, meta);
    return hipMemcpy(p->meta, meta.data(), sizeof(uint32_t) * meta.size(), hipMemcpyHostToDevice);
}

void persistent_layout_meta(int prec, int K, int ipw_req, int& G, int& ipw, int& bpw, std::vector<uint32_t>& meta)
{
    int nb = 0, nBi = 0;
    layout(prec, K, ipw_req, G, ipw, nb, nBi, bpw);
    meta.clear();
    meta.reserve(bpw);
    for (int t = 0; t < ipw; ++t)
        for (int ri = 0; ri < nb; ++ri) {
            const int pr = row_at(ri, nb);
            if (prec == PSF_C_BF16X2) {
                for (int q = pr; q < nb; ++q)
                    meta.push_back(uint32_t(t) << 24 | uint32_t(pr) << 16 | uint32_t(q) << 8 | uint32_t(q));
            } else {
                for (int q = pr; q < nb; q += 2) {
                    const int q2 = q + 1 < nb ? q + 1 : q;
                    meta.push_back(uint32_t(t) << 24 | uint32_t(pr) << 16 | uint32_t(q) << 8 | uint32_t(q2));
                }
            }
        }
}

void persistent_scene_free(psf_scene* s)
{
    PersistentState* p = st(s);
    if (!p) return;
    scene_free(s, p->meta);
    scene_free(s, p->inv_scale);
    scene_free(s, p->xbuf);
    delete p;
    s->persistent = nullptr;
}

namespace {
uint16_t bf16_rne_host(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return static_cast<uint16_t>(b >> 16);
}
float bf16_to_float_host(uint16_t h)
{
    uint32_t b = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
uint16_t f16_host(double v)
{
    const _Float16 h = static_cast<_Float16>(static_cast<float>(v));
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}
}  // namespace

int persistent_upload_c(psf_scene* s, const double* C)
{
    PersistentState* p = st(s);
    if (!p) {
        s->ctx->err = "persistent path not initialised for this precision";
        return PSF_ERR_UNSUPPORTED;
    }
    const int prec = s->d.c_precision;
    const int K = static_cast<int>(s->d.K), Kp = p->Kp, nb = p->nb;
    std::vector<uint16_t> host(size_t(p->G) * p->bpw * 64 * 8, 0);
    std::vector<double> inv(size_t(p->G) * p->ipw, 0.0);
    std::vector<double> T(size_t(Kp) * Kp);   // symmetrised, tiled factor applied
    for (int g = 0; g < p->G; ++g)
        for (int t = 0; t < p->ipw; ++t) {
            const int i = g + p->G * t;
            if (i >= K) continue;
            const double* Ci = C + size_t(i) * K * K;     // C_i(a, b) at b*K + a
            std::fill(T.begin(), T.end(), 0.0);
            double amax = 0.0;
            for (int a = 0; a < K; ++a)
                for (int b = 0; b < K; ++b) {
                    const int pa = a / kTile, qb = b / kTile;
                    if (pa > qb) continue;
                    double v = 0.5 * (Ci[size_t(b) * K + a] + Ci[size_t(a) * K + b]);
                    if (pa < qb) v *= 2.0;
                    T[size_t(a) * Kp + b] = v;
                    amax = std::fmax(amax, std::fabs(v));
                }
            double scale = 1.0;
            if (prec == PSF_C_FP16 && amax > 0.0) scale = std::ldexp(1.0, static_cast<int>(std::floor(std::log2(16384.0 / amax))));
            inv[size_t(g) * p->ipw + t] = 1.0 / (scale * p->wscale);
            // blocks of mode i in metadata order
            int bi = 0;
            for (int ri = 0; ri < nb; ++ri) {
                const int pr = row_at(ri, nb);
                auto emit = [&](int q1, int q2, bool hilo) {
                    const size_t base = ((size_t(g) * p->bpw) + size_t(t) * p->nBi + bi) * 64 * 8;
                    for (int l = 0; l < 64; ++l) {
                        const int r = l & 15, kgp = l >> 4;
                        for (int e = 0; e < 8; ++e) {
                            const int k = 8 * kgp + e;
                            const int row = pr * kTile + r;
                            uint16_t out = 0;
                            if (hilo) {
                                const double v = T[size_t(row) * Kp + q1 * kTile + (k & 15)];
                                const uint16_t hi = bf16_rne_host(static_cast<float>(v));
                                out = k < 16 ? hi : bf16_rne_host(static_cast<float>(v - bf16_to_float_host(hi)));
                            } else {
                                const int q = k < 16 ? q1 : q2;
                                const bool zero = (k >= 16 && q2 == q1);
                                const double v = zero ? 0.0 : T[size_t(row) * Kp + q * kTile + (k & 15)] * scale;
                                out = prec == PSF_C_FP16 ? f16_host(v) : bf16_rne_host(static_cast<float>(v));
                            }
                            host[base + size_t(l) * 8 + e] = out;
                        }
                    }
                    ++bi;
                };
                if (prec == PSF_C_BF16X2) {
                    for (int q = pr; q < nb; ++q) emit(q, q, true);
                } else {
                    for (int q = pr; q < nb; q += 2) emit(q, q + 1 < nb ? q + 1 : q, false);
                }
            }
        }
    // on the context stream, after the steps queued there that read the blocks
    // (a null-stream copy is not ordered against that non-blocking stream)
    hipStream_t cs = s->ctx->stream;
    hipError_t e = hipMemcpyAsync(p->blocks, host.data(), host.size() * 2, hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = hipMemcpyAsync(p->inv_scale, inv.data(), inv.size() * 8, hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = hipStreamSynchronize(cs);
    if (e != hipSuccess) {
        s->ctx->err = std::string("persistent upload: ") + hipGetErrorString(e);
        return PSF_ERR_HIP;
    }
    return PSF_OK;
}

static PArgs make_pargs(psf_scene* s, int nsteps, double* hist, int hist_ld)
{
    PersistentState* p = st(s);
    PArgs a{};
    a.blocks = p->blocks;
    a.meta = p->meta;
    a.inv_scale = p->inv_scale;
    a.xbuf = p->xbuf;
    a.omega = s->dev.omega;
    a.visc = s->dev.visc;
    a.gravity = s->dev.gravity;
    a.status = s->dev.status;
    a.dt = s->dt;
    a.gravity_active = s->gravity_active;
    a.K = static_cast<int>(s->d.K);
    a.Kp = p->Kp;
    a.G = p->G;
    a.ipw = p->ipw;
    a.bpw = p->bpw;
    a.nsteps = nsteps;
    a.tag0 = p->tag;
    a.col1 = p->col1;
    a.wscale = p->wscale;
    a.hist = hist;
    a.hist_ld = hist_ld;
    a.wait_ticks = p->wait_ticks;
    return a;
}

hipError_t launch_persistent_steps(psf_scene* s, int nsteps, hipStream_t stream, double* hist, int hist_ld,
                                   const LaunchEvents* ev)
{
    PersistentState* p = st(s);
    if (!p) return hipErrorNotSupported;
    if (nsteps <= 0) return hipSuccess;
    PArgs a = make_pargs(s, nsteps, hist, hist_ld);
    void* args[] = {&a};
    hipError_t e = ev ? hipExtLaunchKernel(p->fn, dim3(p->G), dim3(64 * kSceneWaves), args, 0, stream, ev->start,
                                           ev->stop, 0)
                      : hipLaunchKernel(p->fn, dim3(p->G), dim3(64 * kSceneWaves), args, 0, stream);
    if (e == hipSuccess) p->tag += static_cast<uint32_t>(nsteps);
    return e;
}

// ---------------------------------------------------------------------------
// Several scenes in one launch.
struct PersistentBatch {
    int max_scenes = 0, max_wgs = 0;
    PArgs* d_table = nullptr;
    uint32_t* d_map = nullptr;
    PArgs* h_table = nullptr;       // pinned staging
    uint32_t* h_map = nullptr;
    int wg_capacity[2] = {0, 0};    // co-resident workgroups of the widest kernel, bf16 / fp16
};

PersistentBatch* persistent_batch_create(int device, int max_scenes, int max_wgs)
{
    PersistentBatch* b = new (std::nothrow) PersistentBatch();
    if (!b) return nullptr;
    b->max_scenes = max_scenes;
    b->max_wgs = max_wgs;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->d_table), sizeof(PArgs) * max_scenes);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_map), sizeof(uint32_t) * max_wgs);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_table), sizeof(PArgs) * max_scenes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_map), sizeof(uint32_t) * max_wgs, hipHostMallocDefault);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    for (int el = 0; el < 2 && e == hipSuccess; ++el) {
        int per_cu = 0;
        const void* fn = el == 0 ? reinterpret_cast<const void*>(&k_persistent_multi<0, kRB>)
                                 : reinterpret_cast<const void*>(&k_persistent_multi<1, kRB>);
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0);
        b->wg_capacity[el] = per_cu * cus;
    }
    if (e != hipSuccess) {
        persistent_batch_destroy(b);
        return nullptr;
    }
    return b;
}

void persistent_batch_destroy(PersistentBatch* b)
{
    if (!b) return;
    if (b->d_table) (void)hipFree(b->d_table);
    if (b->d_map) (void)hipFree(b->d_map);
    if (b->h_table) (void)hipHostFree(b->h_table);
    if (b->h_map) (void)hipHostFree(b->h_map);
    delete b;
}

int persistent_batch_wg_capacity(const PersistentBatch* b, int c_precision)
{
    const int cap = b->wg_capacity[c_precision == PSF_C_FP16 ? 1 : 0];
    return cap < b->max_wgs ? cap : b->max_wgs;
}

hipError_t launch_persistent_batch(PersistentBatch* b, psf_scene* const* sc, const int* nsteps, double* const* hist,
                                   const int* hist_ld, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (n > b->max_scenes) return hipErrorInvalidValue;
    int elem = -1, per_max = 0, wgs = 0;
    for (int j = 0; j < n; ++j) {
        PersistentState* p = st(sc[j]);
        if (!p) return hipErrorNotSupported;
        if (elem >= 0 && p->elem != elem) return hipErrorInvalidValue;
        elem = p->elem;
        per_max = std::max(per_max, (p->bpw + 3) / 4);
        if (nsteps[j] > 0) wgs += p->G;
    }
    if (wgs == 0) return hipSuccess;
    if (wgs > b->max_wgs || wgs > b->wg_capacity[elem]) return hipErrorCooperativeLaunchTooLarge;
    int w = 0;
    for (int j = 0; j < n; ++j) {
        if (nsteps[j] <= 0) continue;
        PersistentState* p = st(sc[j]);
        b->h_table[j] = make_pargs(sc[j], nsteps[j], hist ? hist[j] : nullptr, hist_ld ? hist_ld[j] : 0);
        for (int g = 0; g < p->G; ++g) b->h_map[w++] = static_cast<uint32_t>(j) << 16 | static_cast<uint32_t>(g);
    }
    hipError_t e = hipMemcpyAsync(b->d_table, b->h_table, sizeof(PArgs) * n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_map, b->h_map, sizeof(uint32_t) * wgs, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    const void* fn;
    if (elem == 0)
        fn = per_max <= 4 ? reinterpret_cast<const void*>(&k_persistent_multi<0, 4>)
           : per_max <= 12 ? reinterpret_cast<const void*>(&k_persistent_multi<0, 12>)
                           : reinterpret_cast<const void*>(&k_persistent_multi<0, kRB>);
    else
        fn = per_max <= 4 ? reinterpret_cast<const void*>(&k_persistent_multi<1, 4>)
           : per_max <= 12 ? reinterpret_cast<const void*>(&k_persistent_multi<1, 12>)
                           : reinterpret_cast<const void*>(&k_persistent_multi<1, kRB>);
    const PArgs* tab = b->d_table;
    const uint32_t* map = b->d_map;
    void* args[] = {&tab, &map};
    e = hipLaunchKernel(fn, dim3(wgs), dim3(256), args, 0, stream);
    if (e != hipSuccess) return e;
    for (int j = 0; j < n; ++j)
        if (nsteps[j] > 0) st(sc[j])->tag += static_cast<uint32_t>(nsteps[j]);
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Packing on the device: C_i(a, b) for a, b, i < K taken from a larger FP64
// tensor Cf of K0 modes (C_i(a, b) at Cf[(i * K0 + b) * K0 + a]); the first K
// modes of a basis sorted by eigenvalue give the K-mode tensor.  Same values
// and rounding as persistent_upload_c for BF16 and BF16 hi+lo.
__global__ __launch_bounds__(256) void k_pack_c(const double* __restrict__ Cf, int K0, int K, int G, int ipw, int bpw,
                                                int hilo, const uint32_t* __restrict__ meta, uint4* __restrict__ blocks,
                                                double* __restrict__ inv_scale)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t < static_cast<size_t>(G) * ipw) inv_scale[t] = 1.0;
    const size_t total = static_cast<size_t>(G) * bpw * 64;
    if (t >= total) return;
    const int lane = static_cast<int>(t & 63);
    const size_t gb = t >> 6;
    const int b = static_cast<int>(gb % bpw), g = static_cast<int>(gb / bpw);
    const uint32_t m = meta[b];
    const int tm = static_cast<int>(m >> 24), pr = static_cast<int>((m >> 16) & 255u);
    const int q1 = static_cast<int>((m >> 8) & 255u), q2 = static_cast<int>(m & 255u);
    const int i = g + G * tm;
    const int r = lane & 15, kgp = lane >> 4;
    const int row = pr * kTile + r;
    auto T = [&](int a, int c) -> double {
        if (i >= K || a >= K || c >= K) return 0.0;
        const double* Ci = Cf + static_cast<size_t>(i) * K0 * K0;
        double v = 0.5 * (Ci[static_cast<size_t>(c) * K0 + a] + Ci[static_cast<size_t>(a) * K0 + c]);
        if (a / kTile < c / kTile) v *= 2.0;
        return v;
    };
    uint16_t out[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 8 * kgp + e;
        if (hilo) {
            const double v = T(row, q1 * kTile + (k & 15));
            const uint16_t hi = f32_to_bf16_rne(static_cast<float>(v));
            out[e] = k < 16 ? hi : f32_to_bf16_rne(static_cast<float>(v - static_cast<double>(bf16_to_f32(hi))));
        } else {
            const int q = k < 16 ? q1 : q2;
            out[e] = (k >= 16 && q2 == q1) ? uint16_t(0) : f32_to_bf16_rne(static_cast<float>(T(row, q * kTile + (k & 15))));
        }
    }
    uint4 v;
    v.x = uint32_t(out[0]) | uint32_t(out[1]) << 16;
    v.y = uint32_t(out[2]) | uint32_t(out[3]) << 16;
    v.z = uint32_t(out[4]) | uint32_t(out[5]) << 16;
    v.w = uint32_t(out[6]) | uint32_t(out[7]) << 16;
    blocks[t] = v;
}

hipError_t launch_pack_c(psf_scene* s, const double* Cf, int K0, hipStream_t stream)
{
    PersistentState* p = st(s);
    const int prec = s->d.c_precision;
    if (!p || (prec != PSF_C_BF16 && prec != PSF_C_BF16X2) || K0 < static_cast<int>(s->d.K)) return hipErrorInvalidValue;
    return launch_pack_c_layout(Cf, K0, static_cast<int>(s->d.K), p->G, p->ipw, p->bpw, prec == PSF_C_BF16X2 ? 1 : 0,
                                p->meta, p->blocks, p->inv_scale, stream);
}

hipError_t launch_pack_c_layout(const double* Cf, int K0, int K, int G, int ipw, int bpw, int hilo,
                                const uint32_t* meta, uint4* blocks, double* inv_scale, hipStream_t stream)
{
    const size_t total = std::max(static_cast<size_t>(G) * bpw * 64, static_cast<size_t>(G) * ipw);
    const unsigned grid = static_cast<unsigned>((total + 255) / 256);
    k_pack_c<<<grid, 256, 0, stream>>>(Cf, K0, K, G, ipw, bpw, hilo, meta, blocks, inv_scale);
    return hipGetLastError();
}

int persistent_set_wait_ticks(psf_scene* s, uint64_t ticks)
{
    PersistentState* p = st(s);
    if (!p) return PSF_ERR_UNSUPPORTED;
    p->wait_ticks = ticks ? ticks : kWaitTicks;
    return PSF_OK;
}

}  // namespace psf
