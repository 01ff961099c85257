// HIP kernels of the reduced-space fluid step for gfx950 (MI355X).
//
//   k_contract_plain   partials of y_i = w^T C_i w     psfsolver.cpp:39-44
//   k_update           their sums, then the FP64 energy renormalisation,
//                      viscosity and forcing      psfsolver.cpp:34-38, 45-59
//   k_reconstruct      u = float(Phi w)     psfsolver.cpp:60
//   k_trace            particle pass        psfsolver.cpp:70-232
// (paths relative to the reference repository root)
#include <type_traits>

#include <hip/hip_ext.h>

#include "psf_kernels.hpp"

namespace psf {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Address-space views of pointers that batched launches read from tables in
// memory (the compiler then cannot infer them): global memory, and constant
// memory for data a kernel only reads, so uniform loads become scalar loads.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> as_global(T* p) { return (gptr<T>)(p); }
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p) { return (cptr<T>)(p); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over a 256-thread block in a fixed order (wave butterflies, then waves
// 0..3 in order); every thread receives the result.
__device__ __forceinline__ double block_sum256(double v, double* red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

// ---------------------------------------------------------------------------
// Plain contraction.  C holds K matrices K x K, each column-major (element
// (a, b) of C_i at (i*K + b)*K + a), in FP64 or FP32.  The columns of each
// mode are cut into chunks of kPlainCols: block (i, s) handles columns
// b = 32 s .. 32 s + 31 of C_i, wave w the columns b = 32 s + w + 4 j, and
// writes one FP64 partial sum_b w_b sum_a w_a C_i(a, b) to part[i * S + s];
// k_update adds a mode's S partials in order.  With K * S blocks the launch
// fills the chip (2048 blocks at K = 256), and every lane issues the loads
// of all its columns before the first product, so each wave keeps 8 (FP32)
// or 16 (FP64) vector loads in flight.  Products and the per-column sums
// over a are in T; the column sums enter an FP64 accumulator.
// Measured at K = 256 in FP32: 10.4-12.5 us per launch (5.4-6.5 TB/s) and
// 14.7 us per step with k_update.  Folding the update into the same launch
// (the last block to finish, found through a device-scope counter, sums the
// partials) took 66 us per step: each block's release fence before counting
// writes back its XCD's L2.
constexpr int kPlainCols = 32;

int plain_splits(int K) { return (K + kPlainCols - 1) / kPlainCols; }

template <typename T, int VEC>
__global__ __launch_bounds__(256) void k_contract_plain(const T* __restrict__ C, const double* __restrict__ omega,
                                                        double* __restrict__ part, int K, int S)
{
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sw = reinterpret_cast<T*>(smem_raw);
    __shared__ double red[4];
    for (int k = threadIdx.x; k < K; k += 256) sw[k] = static_cast<T>(omega[k]);
    __syncthreads();
    const int i = blockIdx.x / S, sc = blockIdx.x - i * S;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    constexpr int NJ = kPlainCols / 4;            // columns per wave
    const int b0 = sc * kPlainCols + wv;
    const T* Ci = C + static_cast<size_t>(i) * K * K;
    T colsum[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) colsum[j] = 0;
    for (int a0 = lane * VEC; a0 < K; a0 += kWave * VEC) {
        vec_t cv[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int b = b0 + 4 * j;
            cv[j] = b < K ? *reinterpret_cast<const vec_t*>(Ci + static_cast<size_t>(b) * K + a0) : vec_t{};
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            T t = 0;
#pragma unroll
            for (int e = 0; e < VEC; ++e) t += sw[a0 + e] * cv[j][e];
            colsum[j] += t;
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int b = b0 + 4 * j;
        if (b < K) acc += static_cast<double>(sw[b]) * static_cast<double>(colsum[j]);
    }
    const double sum = block_sum256(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

// FP64 update, one block (psfsolver.cpp:34-59): y_k = sum of the S partials
// part[k * S + s] in order, then
//   e1 = w.w; w += dt*y; e2 = w.w; w *= sqrt(e1/e2); w_k *= exp(nu lambda_k dt);
//   w += dt*g if forcing.
// A zero or non-finite energy sets *status and leaves w unchanged.
// With hist != nullptr the new w is also stored at hist[k * hist_ld + hist_t]
// (one column of a frame window).
// Thread t owns modes t, t + 256, ... (R of them, K <= 256 R) and loads all of
// their inputs before the first reduction, so the kernel pays one memory
// round trip; e1 and e2 are reduced together (one barrier).
template <int R>
__global__ __launch_bounds__(256) void k_update(double* __restrict__ omega, const double* __restrict__ part, int S,
                                                const double* __restrict__ visc, const double* __restrict__ g,
                                                int K, double dt, int gravity_active, int* __restrict__ status,
                                                double* __restrict__ hist, int hist_ld, int hist_t)
{
#pragma clang fp contract(off)
    __shared__ double red[8];
    const int tid = threadIdx.x;
    double w[R], wn[R], vi[R], gi[R];
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = tid + 256 * r;
        w[r] = 0.0; wn[r] = 0.0; vi[r] = 1.0; gi[r] = 0.0;
        if (k < K) {
            // the partials in order, loaded eight at a time (one memory round
            // trip per eight rather than one per partial)
            const double* pk = part + static_cast<size_t>(k) * S;
            double yk = 0.0;
            int q = 0;
            for (; q + 8 <= S; q += 8) {
                double v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = pk[q + e];
#pragma unroll
                for (int e = 0; e < 8; ++e) yk += v[e];
            }
            for (; q < S; ++q) yk += pk[q];
            w[r] = omega[k];
            vi[r] = visc[k];
            gi[r] = g[k];
            wn[r] = w[r] + dt * yk;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        s1 += w[r] * w[r];
        s2 += wn[r] * wn[r];
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    const int lane = tid & (kWave - 1), wv = tid >> 6;
    if (lane == 0) {
        red[wv] = s1;
        red[4 + wv] = s2;
    }
    __syncthreads();
    const double e1 = ((red[0] + red[1]) + red[2]) + red[3];
    const double e2 = ((red[4] + red[5]) + red[6]) + red[7];
    if (!(e2 > 0.0) || !__builtin_isfinite(e2) || !__builtin_isfinite(e1)) {
        if (tid == 0) *status = 1;
        return;
    }
    const double fac = sqrt(e1 / e2);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = tid + 256 * r;
        if (k >= K) break;
        double v = wn[r] * fac;
        v = v * vi[r];
        if (gravity_active) v = v + dt * gi[r];
        omega[k] = v;
        if (hist) hist[static_cast<size_t>(k) * hist_ld + hist_t] = v;
    }
}

// ---------------------------------------------------------------------------
// Reconstruction u_f = float(sum_k Phi[k][f] w_k).  Phi is column-major with a
// zero-padded leading dimension ldF (a multiple of 256).  Every wave owns
// 64*VEC consecutive faces (VEC per lane, one aligned 16-byte load per lane
// and mode) and runs the whole sum for them: an FP64 fma chain over
// k = 0, 1, ..., K-1 from 0.0, rounded to float at the end.  The w_k are
// uniform across the wave and come through scalar loads.  k_reconstruct_mfma
// runs the same chain per face and frame, so both give the same bits.
// FP32 Phi: one face per lane (measured faster than 2 or 4 here), 16 modes in flight
constexpr int kGemvVecF32 = 1;
template <typename T, int VEC>
__global__ __launch_bounds__(256) void k_reconstruct(const T* __restrict__ Phi, size_t ldF, uint32_t F, int K,
                                                     const double* __restrict__ omega, float* __restrict__ u)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t fw = (static_cast<size_t>(blockIdx.x) * 4 + wv) * (kWave * VEC);
    if (fw >= ldF) return;
    const size_t f0 = fw + static_cast<size_t>(lane) * VEC;
    double acc[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j) acc[j] = 0.0;
    const T* base = Phi + f0;
#pragma unroll 16
    for (int k = 0; k < K; ++k) {
        const double w = omega[k];
        if constexpr (sizeof(T) == 4 && VEC == 4) {
            const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(base + static_cast<size_t>(k) * ldF));
            acc[0] = __builtin_fma(static_cast<double>(v.x), w, acc[0]);
            acc[1] = __builtin_fma(static_cast<double>(v.y), w, acc[1]);
            acc[2] = __builtin_fma(static_cast<double>(v.z), w, acc[2]);
            acc[3] = __builtin_fma(static_cast<double>(v.w), w, acc[3]);
        } else if constexpr (sizeof(T) == 4 && VEC == 2) {
            typedef float f32x2 __attribute__((ext_vector_type(2)));
            const f32x2 v = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(base + static_cast<size_t>(k) * ldF));
            acc[0] = __builtin_fma(static_cast<double>(v.x), w, acc[0]);
            acc[1] = __builtin_fma(static_cast<double>(v.y), w, acc[1]);
        } else if constexpr (VEC == 1) {
            const T v = __builtin_nontemporal_load(base + static_cast<size_t>(k) * ldF);
            acc[0] = __builtin_fma(static_cast<double>(v), w, acc[0]);
        } else {
            const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(base + static_cast<size_t>(k) * ldF));
            acc[0] = __builtin_fma(v.x, w, acc[0]);
            acc[1] = __builtin_fma(v.y, w, acc[1]);
        }
    }
#pragma unroll
    for (int j = 0; j < VEC; ++j)
        if (f0 + j < F) u[f0 + j] = static_cast<float>(acc[j]);
}

// ---------------------------------------------------------------------------
// Blocks of a launch are dispatched to the 8 XCDs round-robin (block b to XCD
// b % 8), and each XCD has its own L2.  xcd_block maps launch block b to
// logical block (b % 8) * per + b / 8, so XCD x runs the contiguous logical
// range [x * per, (x + 1) * per): neighbouring particles (stored by voxel)
// and a scene's blocks share one L2, which then holds only their part of u.
// The launch has 8 * per blocks, per = ceil(logical blocks / 8); blocks past
// the logical count exit.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t per) { return (b % kXcds) * per + b / kXcds; }
inline uint32_t xcd_per(uint32_t blocks) { return (blocks + kXcds - 1) / kXcds; }

// Scene of a block in a batched launch: tab[0..n) (n <= 64) by ascending
// block0; one table load per lane and a ballot.
template <typename Tab>
__device__ __forceinline__ int scene_of_block(const Tab* __restrict__ tab, int n, uint32_t blk)
{
    const int lane = threadIdx.x & (kWave - 1);
    const bool le = lane < n && tab[lane].block0 <= blk;
    return __popcll(__ballot(le)) - 1;
}

// ---------------------------------------------------------------------------
// Reconstruction on the FP64 matrix cores: U (faces x frames) = Phi (faces x
// modes) * H (modes x frames), H[k][t] = w_t[k] = hist[k * hist_ld + t] for
// t < W (W = 1 is u = float(Phi w)).  v_mfma_f64_16x16x4f64 adds the four
// products of its k-step to the accumulator as a sequential fma chain
// (checked bit for bit against fma(a3, b3, fma(a2, b2, fma(a1, b1, fma(a0, b0,
// c))))), so a chain of k-steps from 0.0 is the per-face fma chain over
// k = 0..K-1 of k_reconstruct: the same bits, W frames per pass over Phi.
// Modes past K enter as 0 * 0, which leaves every accumulator unchanged.
// A wave owns 64 faces as 4 row tiles of 16, with rows interleaved (row i of
// tile rt is face f0 + 4i + rt), so a lane loads its A operands for all four
// tiles as one aligned 16-byte load of 4 consecutive faces, and stores 4
// consecutive faces of one frame at once (accumulator slot r of lane l holds
// row l/16 + 4r, column l%16).  NT column tiles of 16 frames: W <= 16 * NT.
//
// ACC = float: the same product on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32, twice the FP64 rate on this part) for FP32 Phi
// and w rounded to FP32, with FP32 accumulation: each column is a dot
// product of its own, so a frame's u has the same bits in any window and
// alone (single frames use this kernel with one column).  Its accumulator
// slot r of lane l holds row 4 (l/16) + r (measured), so a lane's 16 values
// of one frame are 16 consecutive faces.
typedef double f64x4 __attribute__((ext_vector_type(4)));

// PAIR: a second scene on the same Phi (K2 <= K modes, its own history and
// fields) rides along: every Phi load feeds both scenes' MFMAs.  Its modes
// k >= K2 enter as 0, so its sums are the MFMAs it would run alone plus
// exact zeros: the same bits.
template <typename T, int NT, typename ACC = double, bool PAIR = false>
__device__ __forceinline__ void mfma_window_body(const T* Phi_, size_t ldF, uint32_t F, int K, const double* hist_,
                                                 int hist_ld, int W, float* uh_, size_t ldU, uint32_t blk,
                                                 int K2 = 0, const double* hist2_ = nullptr, float* uh2_ = nullptr)
{
    static_assert(std::is_same<ACC, double>::value || std::is_same<T, float>::value,
                  "FP32 accumulation takes FP32 Phi");
    constexpr bool kF32 = std::is_same<ACC, float>::value;
    typedef typename std::conditional<kF32, f32x4, f64x4>::type accv;
    const gptr<const T> Phi = as_global(Phi_);
    const gptr<const double> hist = as_global(hist_);
    const gptr<float> uh = as_global(uh_);
    const gptr<const double> hist2 = as_global(hist2_);
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t f0 = (static_cast<size_t>(blk) * 4 + wv) * 64;
    if (f0 >= ldF) return;
    const int li = lane & 15, lk = lane >> 4;
    const gptr<const T> arow = Phi + f0 + 4 * li;
    constexpr int NS = PAIR ? 2 : 1;   // scenes
    accv acc[NS][4][NT];
#pragma unroll
    for (int sc = 0; sc < NS; ++sc)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) acc[sc][rt][ct] = accv{0, 0, 0, 0};
    // frames of the lane's B column: t = 16 ct + li; columns t >= W load a
    // valid element (t clamped) and multiply by 0 instead
    bool tv[NT];
    int tc[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
        tv[ct] = 16 * ct + li < W;
        tc[ct] = tv[ct] ? 16 * ct + li : W - 1;
    }
    // operands as loaded (FP32 Phi is widened at use, and B columns past W
    // are zeroed at use, so no load is waited for before its step)
    struct RawA {
        typename std::conditional<sizeof(T) == 4, f32x4, f64x2>::type v[sizeof(T) == 4 ? 1 : 2];
    };
    auto load_a = [&](int k, RawA& r) {
        if constexpr (sizeof(T) == 4) {
            r.v[0] = __builtin_nontemporal_load(reinterpret_cast<gptr<const f32x4>>(arow + static_cast<size_t>(k) * ldF));
        } else {
            const gptr<const f64x2> q = reinterpret_cast<gptr<const f64x2>>(arow + static_cast<size_t>(k) * ldF);
            r.v[0] = __builtin_nontemporal_load(q);
            r.v[1] = __builtin_nontemporal_load(q + 1);
        }
    };
    // B values of both scenes; the second's modes past K2 load mode K2 - 1
    // (in bounds) and enter as 0
    auto load_b = [&](int k, double (&b)[NS][NT]) {
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) b[0][ct] = hist[static_cast<size_t>(k) * hist_ld + tc[ct]];
        if constexpr (PAIR) {
            const int k2 = k < K2 ? k : K2 - 1;
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) b[1][ct] = hist2[static_cast<size_t>(k2) * hist_ld + tc[ct]];
        }
    };
    auto step = [&](const RawA& r, const double (&braw)[NS][NT], int k) {
        ACC a[4];
        if constexpr (sizeof(T) == 4) {
            a[0] = r.v[0].x; a[1] = r.v[0].y; a[2] = r.v[0].z; a[3] = r.v[0].w;
        } else {
            a[0] = r.v[0].x; a[1] = r.v[0].y; a[2] = r.v[1].x; a[3] = r.v[1].y;
        }
#pragma unroll
        for (int sc = 0; sc < NS; ++sc) {
            const bool kv = sc == 0 || k < K2;
            ACC b[NT];
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) b[ct] = (tv[ct] && kv) ? static_cast<ACC>(braw[sc][ct]) : ACC(0);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) {
                    if constexpr (kF32)
                        acc[sc][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], b[ct], acc[sc][rt][ct], 0, 0, 0);
                    else
                        acc[sc][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[sc][rt][ct], 0, 0, 0);
                }
        }
    };
    // full k-steps (every k < K), branch-free, loads kPipe steps ahead: the
    // queue slot of step k4 is refilled with step k4 + kPipe (clamped to the
    // last full step; the extra loads are in bounds and unused)
    constexpr int kPipe = PAIR ? 3 : 4;   // (6, 8 and 12 measured no faster with FP32 accumulation)
    const int K4 = K >> 2;
    if (K4 > 0) {
        RawA qa[kPipe];
        double qb[kPipe][NS][NT];
#pragma unroll
        for (int st = 0; st < kPipe; ++st) {
            const int k = 4 * (st < K4 ? st : K4 - 1) + lk;
            load_a(k, qa[st]);
            load_b(k, qb[st]);
        }
        int k4 = 0;
        for (; k4 + kPipe <= K4; k4 += kPipe) {
#pragma unroll
            for (int st = 0; st < kPipe; ++st) {
                step(qa[st], qb[st], 4 * (k4 + st) + lk);
                const int nx = k4 + st + kPipe;
                const int k = 4 * (nx < K4 ? nx : K4 - 1) + lk;
                load_a(k, qa[st]);
                load_b(k, qb[st]);
                // keep the refill here: the scheduler would otherwise sink
                // the loads to the loop's end and wait for them there
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int st = 0; st < kPipe; ++st)
            if (st < K4 - k4) step(qa[st], qb[st], 4 * (k4 + st) + lk);
    }
    // the last, partial k-step: modes past K enter as 0 * 0
    if (K & 3) {
        const int k = 4 * K4 + lk;
        RawA ra{};
        double rb[NS][NT];
#pragma unroll
        for (int sc = 0; sc < NS; ++sc)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) rb[sc][ct] = 0.0;
        if (k < K) {
            load_a(k, ra);
            load_b(k, rb);
        }
        step(ra, rb, k);
    }
    // The stores go through a wave-private LDS tile of 16 frames x 64 faces:
    // a lane's values of one frame are 16-byte pieces 64 bytes apart, and
    // stores of that shape ran at about 1 TB/s (a third of the launch at
    // W = 16).  Read back by rows, each store instruction writes 256
    // contiguous bytes (a wave's 64 faces) of each of four frames.
    constexpr int kTileLd = 68;   // floats per tile row (padded)
    __shared__ __attribute__((aligned(16))) float tiles[4][16 * kTileLd];
    float* const tile = tiles[wv];
#pragma unroll
    for (int sc = 0; sc < NS; ++sc)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
        if (16 * ct >= W) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // local face of slot r in tile 0 (tile rt adds rt)
            const int fl = kF32 ? 16 * lk + 4 * r : 4 * (lk + 4 * r);
            const f32x4 o = {static_cast<float>(acc[sc][0][ct][r]), static_cast<float>(acc[sc][1][ct][r]),
                             static_cast<float>(acc[sc][2][ct][r]), static_cast<float>(acc[sc][3][ct][r])};
            *reinterpret_cast<f32x4*>(tile + li * kTileLd + fl) = o;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int row = lk + 4 * s4, t = 16 * ct + row;
            const f32x4 o = *reinterpret_cast<const f32x4*>(tile + row * kTileLd + 4 * li);
            if (t >= W) continue;
            const gptr<float> ut = (sc == 0 ? uh : as_global(uh2_)) + static_cast<size_t>(t) * ldU;
            const size_t f = f0 + 4 * li;
            if (f + 3 < F) {
                __builtin_nontemporal_store(o, reinterpret_cast<gptr<f32x4>>(ut + f));
            } else {
                if (f < F) ut[f] = o.x;
                if (f + 1 < F) ut[f + 1] = o.y;
                if (f + 2 < F) ut[f + 2] = o.z;
            }
        }
        // the next tile's writes follow these reads
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

template <typename T, int NT, typename ACC = double>
__global__ __launch_bounds__(256) void k_reconstruct_mfma(const T* __restrict__ Phi, size_t ldF, uint32_t F, int K,
                                                          const double* __restrict__ hist, int hist_ld, int W,
                                                          float* __restrict__ uh, size_t ldU)
{
    mfma_window_body<T, NT, ACC>(Phi, ldF, F, K, hist, hist_ld, W, uh, ldU, blockIdx.x);
}

template <int NT, typename ACC>
__global__ __launch_bounds__(256) void k_reconstruct_mfma_multi(const WindowScene* __restrict__ tab, int n, int W)
{
    const WindowScene& w = tab[scene_of_block(tab, n, blockIdx.x)];
    mfma_window_body<float, NT, ACC>(static_cast<const float*>(w.Phi), w.ldF, w.F, w.K, w.hist, W, W, w.uh, w.ldF,
                                     blockIdx.x - w.block0);
}

// Entries with a second scene (K2 > 0) on the same Phi: one pass over Phi for
// both (FP32 sums)
template <int NT>
__global__ __launch_bounds__(256) void k_reconstruct_mfma_pairs(const WindowScene* __restrict__ tab, int n, int W)
{
    const WindowScene& w = tab[scene_of_block(tab, n, blockIdx.x)];
    if (w.K2 > 0)
        mfma_window_body<float, NT, float, true>(static_cast<const float*>(w.Phi), w.ldF, w.F, w.K, w.hist, W, W, w.uh,
                                                 w.ldF, blockIdx.x - w.block0, w.K2, w.hist2, w.uh2);
    else
        mfma_window_body<float, NT, float>(static_cast<const float*>(w.Phi), w.ldF, w.F, w.K, w.hist, W, W, w.uh,
                                           w.ldF, blockIdx.x - w.block0);
}

// out[k] = hist[k * ld + t]: one column of a window history
__global__ __launch_bounds__(256) void k_gather_column(const double* __restrict__ hist, int ld, int t, int K,
                                                       double* __restrict__ out)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < K) out[k] = hist[static_cast<size_t>(k) * ld + t];
}

// ---------------------------------------------------------------------------
// Particle pass, one thread per particle.  The arithmetic restates
// psfsolver.cpp:78-222 operation by operation with contraction off, so the
// cell lookups and the float results match the CPU path bit for bit.

// face sign of voxel (x,y,z), face j: the closed form of the sign bit string
// (decmesh3d.cpp:1383-1399): a face is positive iff the voxel created it,
// i.e. {z==0, 1, y==0, 1, x==0, 1}.
__device__ __forceinline__ float face_sign(int j, int x, int y, int z)
{
    bool pos = (j & 1) ? true : (j == 0 ? z == 0 : (j == 2 ? y == 0 : x == 0));
    return pos ? 1.0f : -1.0f;
}

__device__ __forceinline__ float mixf(float x, float y, float a)
{
#pragma clang fp contract(off)
    return x * (1.0f - a) + y * a;   // glm::mix
}

// v < lo ? lo : (v > hi ? hi : v) for lo <= hi, written as the median of
// (v, lo, hi) in the form the backend selects as one v_med3_i32
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return max(min(v, lo), min(max(v, lo), hi)); }

// The three cell indices of one axis as the CPU path computes them
// (psfsolver.cpp:80-100): O = int(a / h), M = int((a - h/2) / h),
// P = int((a + h/2) / h), each quotient rounded to double before the
// truncation.  From rb = RN(1/h), q = RN(a * rb) is within 3 * 2^-53 |q| of
// a / h, and q -+ 1/2 are as close to the other two quotients; so when q lies
// farther than 2^-40 max(|q|, 1/2) from every multiple of 1/2, no integer
// separates any computed quotient from the CPU's and they truncate alike.
// Otherwise (particles within ~1e-12 cells of a voxel or half-voxel plane,
// or NaN) the three IEEE divisions decide.
__device__ __forceinline__ void cell_indices(double a, double h, double hh, double rb, int& O, int& M, int& P)
{
    const double q = a * rb, q2 = q + q;
    if (__builtin_fabs(q2 - __builtin_rint(q2)) > 0x1p-39 * __builtin_fmax(__builtin_fabs(q2), 1.0)) {
        O = static_cast<int>(q);
        M = static_cast<int>(q - 0.5);
        P = static_cast<int>(q + 0.5);
    } else {
        O = static_cast<int>(a / h);
        M = static_cast<int>((a - hh) / h);
        P = static_cast<int>((a + hh) / h);
    }
}

// The same three indices when h = 2^-k, in float; equal to cell_indices'
// once clamped to [0, n - 1] (n <= 2048), which every caller does.
// q = a / h = a * 2^k is exact in float and in double, and the CPU path's
// (a -+ h/2) / h is q -+ 1/2 (exactly, or where a is too small or too large
// for that, with the same clamped index).  For 1/2 <= q < 2^23, q - 1/2 is
// exact in float (a multiple of q's ulp <= 1/2, below q), and q + 1/2 either
// is exact or rounds within [2^j, 2^j + 1/2] after crossing a power of two
// 2^j >= 1, so truncation agrees.  Below 1/2 every index clamps to 0 (q + 1/2
// alone could round up to 1 in float, hence the select); from 2^23 up all
// three clamp to n - 1; NaN converts to 0 in both.
__device__ __forceinline__ void cell_indices_pow2(float a, float rb, int& O, int& M, int& P)
{
#pragma clang fp contract(off)
    const float q = a * rb;
    O = static_cast<int>(q);
    M = static_cast<int>(q - 0.5f);
    P = q < 0.5f ? 0 : static_cast<int>(q + 0.5f);
}

// two consecutive face fluxes u[i], u[i + 1] in one 4-byte-aligned load (i + 1 < F
// always holds for the faces paired below)
typedef float FluxPair __attribute__((ext_vector_type(2))) __attribute__((aligned(4)));

// The three respawn draws in [0, 1024) of (seed, step, particle): one
// SplitMix64 output (Steele, Lea and Flood 2014: a Weyl increment, then the
// variant-13 finaliser) of a key mixing the three counters, cut into 10-bit
// fields, draw k = bits 10k .. 10k + 9.
struct RespawnDraws {
    uint32_t r0, r1, r2;
};
__device__ __forceinline__ RespawnDraws respawn_draws(uint32_t seed, uint32_t step, uint32_t particle)
{
    uint64_t z = (static_cast<uint64_t>(seed) << 32) ^ (static_cast<uint64_t>(step) * 0x9E3779B97F4A7C15ull) ^
                 (static_cast<uint64_t>(particle) * 0xD1B54A32D192ED03ull);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    const uint32_t lo = static_cast<uint32_t>(z ^ (z >> 31));
    return {lo & 1023u, (lo >> 10) & 1023u, (lo >> 20) & 1023u};
}

// A particle's cell indices on each axis (O: its voxel, M/P: the voxels
// below/above the half-voxel-shifted point), clamped to the grid.
struct TraceCell {
    int xO, xM, xP, yO, yM, yP, zO, zM, zP;
};

// Pow2: the voxel size is a power of two (cell_indices_pow2)
template <bool Pow2>
__device__ __forceinline__ TraceCell trace_cell(const TraceParams& P, const float4 p)
{
#pragma clang fp contract(off)
    const int nx = static_cast<int>(P.nx), ny = static_cast<int>(P.ny), nz = static_cast<int>(P.nz);
    const float sx = (P.aabb_min[0] + P.ext[0]) + p.x;
    const float sy = (P.aabb_min[1] + P.ext[1]) + p.y;
    const float sz = (P.aabb_min[2] + P.ext[2]) + p.z;
    TraceCell c;
    if constexpr (Pow2) {
        cell_indices_pow2(sx, P.inv_h, c.xO, c.xM, c.xP);
        cell_indices_pow2(sy, P.inv_h, c.yO, c.yM, c.yP);
        cell_indices_pow2(sz, P.inv_h, c.zO, c.zM, c.zP);
    } else {
        const double res = P.res, hres = P.res / 2, rr = P.inv_res;
        cell_indices(static_cast<double>(sx), res, hres, rr, c.xO, c.xM, c.xP);
        cell_indices(static_cast<double>(sy), res, hres, rr, c.yO, c.yM, c.yP);
        cell_indices(static_cast<double>(sz), res, hres, rr, c.zO, c.zM, c.zP);
    }
    c.xO = clampi(c.xO, 0, nx - 1); c.yO = clampi(c.yO, 0, ny - 1); c.zO = clampi(c.zO, 0, nz - 1);
    c.xM = clampi(c.xM, 0, nx - 1); c.xP = clampi(c.xP, 0, nx - 1);
    c.yM = clampi(c.yM, 0, ny - 1); c.yP = clampi(c.yP, 0, ny - 1);
    c.zM = clampi(c.zM, 0, nz - 1); c.zP = clampi(c.zP, 0, nz - 1);
    return c;
}

// The 12 flux pairs a particle pass loads: x faces of voxel xO at rows
// (yM|yP, zM|zP), and y and z faces of columns (xM, xM+1) at their rows.
struct TraceFlux {
    FluxPair x[4], y[4], z[4];
};

__device__ __forceinline__ uint32_t cell_dy(const TraceCell& c) { return static_cast<uint32_t>(c.yP - c.yM); }
__device__ __forceinline__ uint32_t cell_dz(const TraceCell& c) { return static_cast<uint32_t>(c.zP - c.zM); }

// Byte offsets of the first face of each y and z pair (the x pairs are
// derived in trace_fetch).  Face indices (decmesh3d.cpp:17-30): Z faces, then
// Y, then X, each block z-major.  Strides stay below 2^24 and indices below
// 2^30, so the products run on the 24-bit multiplier and byte offsets from
// the uniform base fit 32 bits (4 F < 2^32); neighbouring faces are reached
// by adding strides (the M/P pairs differ by 0, 1 or 2 cells).
__device__ __forceinline__ void yz_offsets(const TraceParams& P, const TraceCell& c, uint32_t iy[4], uint32_t iz[4])
{
    const uint32_t sZy = P.nx, sZz = P.nx * P.ny;
    const uint32_t sYy = P.nx, sYz = (P.ny + 1) * P.nx;
    iy[0] = (__umul24(c.zM, sYz) + __umul24(c.yO, sYy) + c.xM + P.nZF) << 2;
    iy[1] = iy[0] + (sYy << 2);
    iy[2] = iy[0] + (__umul24(cell_dz(c), sYz) << 2);
    iy[3] = iy[2] + (sYy << 2);
    iz[0] = (__umul24(c.zO, sZz) + __umul24(c.yM, sZy) + c.xM) << 2;
    iz[1] = iz[0] + (__umul24(cell_dy(c), sZy) << 2);
    iz[2] = iz[0] + (sZz << 2);
    iz[3] = iz[1] + (sZz << 2);
}

// The flux pairs of cell c in field u.  Faces of neighbouring x are adjacent
// in every block of the layout, so each (x, x+1) couple comes in one load:
// the x faces of voxel xO, and the y and z faces of columns xM and xP
// (xP = xM + 1 but at clamps).  Only loads are issued here; the values are
// first used in trace_advance.
__device__ __forceinline__ TraceFlux trace_fetch(const TraceParams& P, const gptr<const float> u, const TraceCell& c)
{
    const uint32_t sXy = P.nx + 1, sXz = P.ny * (P.nx + 1);
    const gptr<const char> ub = reinterpret_cast<gptr<const char>>(u);
    auto pair = [&](uint32_t b) { return *reinterpret_cast<gptr<const FluxPair>>(ub + b); };
    TraceFlux f;
    const uint32_t bx = (__umul24(c.zM, sXz) + __umul24(c.yM, sXy) + c.xO + P.nZF + P.nYF) << 2;
    const uint32_t dyx = __umul24(cell_dy(c), sXy) << 2, dzx = __umul24(cell_dz(c), sXz) << 2;
    f.x[0] = pair(bx); f.x[1] = pair(bx + dyx); f.x[2] = pair(bx + dzx); f.x[3] = pair(bx + dyx + dzx);
    uint32_t iy[4], iz[4];
    yz_offsets(P, c, iy, iz);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f.y[k] = pair(iy[k]);
        f.z[k] = pair(iz[k]);
    }
    return f;
}

// One step of a live particle (psfsolver.cpp:102-215) from its cell and flux
// pairs (u: the field the pairs came from, for the rare patch below)
__device__ __forceinline__ float4 trace_advance(const TraceParams& P, const gptr<const float> u, float4 p,
                                                const TraceCell& c, const TraceFlux& f)
{
#pragma clang fp contract(off)
    // the x+ neighbour from each y and z pair: dx is 1 away from the walls,
    // 0 at a clamp (the pair's first face again); a dx of any other value is
    // patched with its own loads.  Waves without a clamped particle skip the
    // patch.
    const int dx = c.xP - c.xM;
    float r3y = f.y[0].y, r4y = f.y[1].y, r7y = f.y[2].y, r8y = f.y[3].y;
    float r3z = f.z[0].y, r4z = f.z[1].y, r7z = f.z[2].y, r8z = f.z[3].y;
    if (__builtin_expect(dx != 1, 0)) {
        if (dx == 0) {
            r3y = f.y[0].x; r4y = f.y[1].x; r7y = f.y[2].x; r8y = f.y[3].x;
            r3z = f.z[0].x; r4z = f.z[1].x; r7z = f.z[2].x; r8z = f.z[3].x;
        } else {
            const gptr<const char> ub = reinterpret_cast<gptr<const char>>(u);
            auto one = [&](uint32_t b) { return *reinterpret_cast<gptr<const float>>(ub + b); };
            uint32_t iy[4], iz[4];
            yz_offsets(P, c, iy, iz);
            const uint32_t db = static_cast<uint32_t>(dx) << 2;
            r3y = one(iy[0] + db); r4y = one(iy[1] + db); r7y = one(iy[2] + db); r8y = one(iy[3] + db);
            r3z = one(iz[0] + db); r4z = one(iz[1] + db); r7z = one(iz[2] + db); r8z = one(iz[3] + db);
        }
    }
    const float s0 = face_sign(0, 0, 0, c.zO), s1 = face_sign(1, 0, 0, c.zO);
    const float s2 = face_sign(2, 0, c.yO, 0), s3 = face_sign(3, 0, c.yO, 0);
    const float s4 = face_sign(4, c.xO, 0, 0), s5 = face_sign(5, c.xO, 0, 0);
    // Face centres: min + h * i, plus h/2 on the axes the face spans.  On the
    // face's own axis the CPU adds 0.5f * 0.0f; min + h * i is never -0
    // (h > 0, i >= 0, and -0 + +0 = +0) and finite for a valid AABB, so adding
    // +0 leaves it unchanged and is skipped.
    const float hf = P.hf;
    const float cx0 = P.aabb_min[0] + hf * static_cast<float>(c.xO);
    const float cx1 = (P.aabb_min[1] + hf * static_cast<float>(c.yM)) + 0.5f * hf;
    const float cx2 = (P.aabb_min[2] + hf * static_cast<float>(c.zM)) + 0.5f * hf;
    const float cy0 = (P.aabb_min[0] + hf * static_cast<float>(c.xM)) + 0.5f * hf;
    const float cy1 = P.aabb_min[1] + hf * static_cast<float>(c.yO);
    const float cy2 = (P.aabb_min[2] + hf * static_cast<float>(c.zM)) + 0.5f * hf;
    const float cz0 = (P.aabb_min[0] + hf * static_cast<float>(c.xM)) + 0.5f * hf;
    const float cz1 = (P.aabb_min[1] + hf * static_cast<float>(c.yM)) + 0.5f * hf;
    const float cz2 = P.aabb_min[2] + hf * static_cast<float>(c.zO);
    const float inv = P.inv_h;
    const float nx0 = inv * (p.x - cx0), nx1 = inv * (p.y - cx1), nx2 = inv * (p.z - cx2);
    const float ny0 = inv * (p.x - cy0), ny1 = inv * (p.y - cy1), ny2 = inv * (p.z - cy2);
    const float nz0 = inv * (p.x - cz0), nz1 = inv * (p.y - cz1), nz2 = inv * (p.z - cz2);
    // The CPU path mixes the signed face velocities vel_k = s * u_k
    // (psfsolver.cpp:160-206): first pairwise along one axis, then the pairs,
    // then the halves.  In the first level glm::mix(s a, s' b, t) =
    // (s a)(1 - t) + (s' b) t, and (s a) w = a (s w) exactly for s = +-1
    // (round to nearest commutes with negation), so the signs ride on the
    // four weights instead of the eight fluxes.
    float vx, vy, vz;
    {
        // vel1,2,5,6 = -s4 * (left faces), vel3,4,7,8 = s5 * (right faces)
        const float w = 1.0f - nx1, la = -s4 * w, lb = -s4 * nx1, ra = s5 * w, rb = s5 * nx1;
        const float a0 = f.x[0].x * la + f.x[1].x * lb, a1 = f.x[0].y * ra + f.x[1].y * rb;
        const float a2 = f.x[2].x * la + f.x[3].x * lb, a3 = f.x[2].y * ra + f.x[3].y * rb;
        vx = mixf(mixf(a0, a1, nx0), mixf(a2, a3, nx0), nx2);
    }
    {
        // odd vel = -s2 * (bottom faces), even vel = s3 * (top faces)
        const float wa = -s2 * (1.0f - ny1), wb = s3 * ny1;
        const float a0 = f.y[0].x * wa + f.y[1].x * wb, a1 = r3y * wa + r4y * wb;
        const float a2 = f.y[2].x * wa + f.y[3].x * wb, a3 = r7y * wa + r8y * wb;
        vy = mixf(mixf(a0, a1, ny0), mixf(a2, a3, ny0), ny2);
    }
    {
        // vel1..4 = s0 * (front faces), vel5..8 = -s1 * (back faces)
        const float w = 1.0f - nz1, fa = s0 * w, fb = s0 * nz1, ba = -s1 * w, bb = -s1 * nz1;
        const float a0 = f.z[0].x * fa + f.z[1].x * fb, a1 = r3z * fa + r4z * fb;
        const float a2 = f.z[2].x * ba + f.z[3].x * bb, a3 = r7z * ba + r8z * bb;
        vz = mixf(mixf(a0, a1, nz0), mixf(a2, a3, nz0), nz2);
    }
    // glm::clamp(v, lo, hi) = min(max(v, lo), hi), for lo < hi the median
    // of (v, lo, hi): one v_med3_f32.  (The two differ only for a NaN v,
    // which the CPU keeps, and for v = -0 on a bound of 0.)
    auto clampf = [](float v, float lo, float hi) { return __builtin_amdgcn_fmed3f(v, lo, hi); };
    p.x = clampf(p.x + P.dtf * vx, P.aabb_min[0], P.aabb_max[0]);
    p.y = clampf(p.y + P.dtf * vy, P.aabb_min[1], P.aabb_max[1]);
    p.z = clampf(p.z + P.dtf * vz, P.aabb_min[2], P.aabb_max[2]);
    p.w = p.w - 1.0f;
    return p;
}

// respawn of an expired particle (psfsolver.cpp:219-222, with the build's
// counter-based draws); the draws follow the caller's particle index who
__device__ __forceinline__ float4 trace_respawn(const TraceParams& P, uint32_t who)
{
    const RespawnDraws d = respawn_draws(P.seed, P.frame, who);
    const uint32_t r0 = d.r0, r1 = d.r1, r2 = d.r2;
    float4 p;
    p.x = static_cast<float>((r0 / 1024.0 - 0.5) * 0.25);
    p.y = P.aabb_min[1] + 0.2f;
    p.z = static_cast<float>(-static_cast<double>(P.center_z) + (r1 / 1024.0 - 0.5) * 0.25);
    p.w = static_cast<float>(P.lifetime * (r2 / 1024.0));
    return p;
}

// one particle record through one frame; id is its storage slot (cells),
// who() its caller index (called only for a respawn)
template <bool Pow2, typename WhoFn>
__device__ __forceinline__ float4 trace_one(const TraceParams& P, const gptr<const float> u, float4 p, WhoFn who,
                                            uint32_t id, int32_t* __restrict__ cells)
{
    if (p.w > 0.0f) {
        const TraceCell c = trace_cell<Pow2>(P, p);
        if (cells) {
            int32_t* o = cells + 12ull * id;
            o[0] = c.xO; o[1] = c.yO; o[2] = c.zO; o[3] = c.xM; o[4] = c.xP; o[5] = c.yM; o[6] = c.yP;
            o[7] = c.zM; o[8] = c.zP; o[9] = 0; o[10] = 0; o[11] = 0;
        }
        return trace_advance(P, u, p, c, trace_fetch(P, u, c));
    }
    if (cells) {
        int32_t* o = cells + 12ull * id;
#pragma unroll
        for (int k = 0; k < 12; ++k) o[k] = -1;
    }
    return trace_respawn(P, who());
}

// One counter increment per particle, with equal cells of neighbouring lanes
// merged: the first lane of a run of equal cells adds the run's length.  A
// wave's atomic costs about one L2 request per distinct line plus one per
// repeated address, so runs (particles stored by voxel, piled up at a wall)
// become one request.  Callers run it with every lane of the wave that holds
// a particle; the active lanes are a prefix of the wave (consecutive ids,
// the tail cut off by the particle count).
__device__ __forceinline__ void density_add_key(uint32_t* counts, uint32_t key)
{
    const int lane = static_cast<int>(threadIdx.x & (kWave - 1));
    const uint64_t active = __ballot(1);
    const uint32_t prev = __shfl_up(key, 1, kWave);
    const bool head = lane == 0 || prev != key;
    const uint64_t ends = __ballot(head) | ~active;   // lanes where a run starts, or past the last particle
    if (!head) return;
    const uint64_t above = lane == kWave - 1 ? 0ull : ends & (~0ull << (lane + 1));
    const int next = above ? __builtin_ctzll(above) : kWave;
    __hip_atomic_fetch_add(counts + key, static_cast<uint32_t>(next - lane), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void density_add(const DensityGrid& D, float x, float y, float z)
{
    density_add_key(D.counts, density_cell(D, x, y, z));
}

template <bool Pow2>
__global__ __launch_bounds__(256) void k_trace(TraceParams P, const float* __restrict__ u, float4* __restrict__ parts,
                                               const uint32_t* __restrict__ pid, uint32_t n, int32_t* __restrict__ cells,
                                               DensityGrid D, uint32_t per)
{
    const uint32_t id = xcd_block(blockIdx.x, per) * 256u + threadIdx.x;
    if (id >= n) return;
    // particle records stream through once: nontemporal, so u stays in L2
    f32x4* rec = reinterpret_cast<f32x4*>(parts) + id;
    const f32x4 v = __builtin_nontemporal_load(rec);
    const gptr<const uint32_t> gp = as_global(pid);
    const float4 p = trace_one<Pow2>(P, as_global(u), make_float4(v.x, v.y, v.z, v.w), [&] { return gp[id]; }, id,
                                     cells);
    const f32x4 o = {p.x, p.y, p.z, p.w};
    __builtin_nontemporal_store(o, rec);
    if (D.counts) density_add(D, p.x, p.y, p.z);
}

// Several frames' particle passes in one launch: each thread carries its
// particle through frames t = 0..nframes-1 with u_t = u0 + t * ldU and the
// respawn frame P.frame + t, and stores it once (the same trace_one per
// frame, so the bits of nframes single passes).  With D.counts set the
// positions after the last frame are counted.
// 8 waves per SIMD (at most 64 VGPRs; both instances fit without spilling):
// the windowed frame loop, where the next window's reconstruction and steps
// share the CUs, ran 5% faster than at 70 VGPRs.  Loading the next frame's
// fluxes one frame early (with the current cell) measured 8% slower: it
// needs 96 VGPRs (5 waves per SIMD).  Two LDS-staging variants were
// measured and gave no gain, with the same bits: per frame, each block
// reducing its particles' box of cells and staging the faces in it (two
// barriers per frame) took 15.5 us per frame against 12; staging the box's
// faces of all 8 frames once per launch (36 KiB per block, 4 blocks per CU)
// took 12 us.  With the address unit relieved the time stays: the pass is
// bound by its ~220 vector instructions per particle-frame (about half the
// issue cycles) and the dependent chain from one frame's position to the
// next frame's gathers.
template <bool Pow2>
__global__ __launch_bounds__(256, 8) void k_trace_frames(TraceParams P, const float* __restrict__ u0, size_t ldU,
                                                      uint32_t nframes, float4* __restrict__ parts,
                                                      const uint32_t* __restrict__ pid, uint32_t n, DensityGrid D,
                                                      uint32_t per)
{
    const uint32_t id = xcd_block(blockIdx.x, per) * 256u + threadIdx.x;
    if (id >= n) return;
    f32x4* rec = reinterpret_cast<f32x4*>(parts) + id;
    const f32x4 v = __builtin_nontemporal_load(rec);
    float4 p = make_float4(v.x, v.y, v.z, v.w);
    const gptr<const float> u = as_global(u0);
    // the caller index, loaded once: a respawn then waits for no load
    const uint32_t who = as_global(pid)[id];
    for (uint32_t t = 0; t < nframes; ++t) {
        p = trace_one<Pow2>(P, u + static_cast<size_t>(t) * ldU, p, [who] { return who; }, id, nullptr);
        P.frame++;
    }
    const f32x4 o = {p.x, p.y, p.z, p.w};
    __builtin_nontemporal_store(o, rec);
    if (D.counts) density_add(D, p.x, p.y, p.z);
}

// Density of the stored particles (the histogram pass of the GPU solver's
// particle drawing, psfsolvergpu.cpp:893-899).  Particles often gather in
// few cells (they respawn on a small emitter sheet and live at most
// `lifetime` frames), and global atomics on one address serialise in its L2
// channel.  So each block loads a chunk of kDensChunk particles, counts their
// cells into an LDS hash table (open addressing, at most kDensChunk keys in
// twice as many slots) and then:
//   - where that takes fewer global atomics than merging runs of equal cells
//     in neighbouring lanes (distinct cells < half the runs: a hash-order
//     atomic touches its own line, storage-order ones share lines), adds
//     each distinct cell's count with one global atomic;
//   - otherwise adds per particle with run-merged atomics in storage order.
constexpr int kDensPer = 8;                       // particles per thread
constexpr int kDensChunk = 256 * kDensPer;
constexpr int kDensSlots = 2 * kDensChunk;
constexpr uint32_t kDensEmpty = 0xffffffffu;      // cell indexes are below 2^30

__global__ __launch_bounds__(256) void k_density_count(DensityGrid D, const float4* __restrict__ parts, uint32_t n)
{
    __shared__ uint32_t key[kDensSlots];
    __shared__ uint32_t cnt[kDensSlots];
    __shared__ uint32_t distinct, runs;
    for (int s = threadIdx.x; s < kDensSlots; s += 256) {
        key[s] = kDensEmpty;
        cnt[s] = 0u;
    }
    if (threadIdx.x == 0) {
        distinct = 0u;
        runs = 0u;
    }
    const uint32_t base = blockIdx.x * static_cast<uint32_t>(kDensChunk) + threadIdx.x;
    uint32_t k[kDensPer];
#pragma unroll
    for (int j = 0; j < kDensPer; ++j) {
        const uint32_t id = base + static_cast<uint32_t>(j * 256);
        f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
        if (id < n) v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(parts) + id);
        k[j] = id < n ? density_cell(D, v.x, v.y, v.z) : kDensEmpty;
    }
    __syncthreads();
    uint32_t my_runs = 0;
    const int lane = static_cast<int>(threadIdx.x & (kWave - 1));
#pragma unroll
    for (int j = 0; j < kDensPer; ++j) {
        const uint32_t prev = __shfl_up(k[j], 1, kWave);
        my_runs += static_cast<uint32_t>(__builtin_popcountll(__ballot(k[j] != kDensEmpty && (lane == 0 || prev != k[j]))));
    }
    if (lane == 0) atomicAdd(&runs, my_runs);
#pragma unroll
    for (int j = 0; j < kDensPer; ++j) {
        if (k[j] == kDensEmpty) continue;
        uint32_t h = (k[j] * 2654435761u) >> (32 - 12);
        static_assert(kDensSlots == 1 << 12, "hash width");
        for (int probe = 0; probe < kDensSlots; ++probe) {   // a free or matching slot exists
            const uint32_t prev = atomicCAS(&key[h], kDensEmpty, k[j]);
            if (prev == kDensEmpty) atomicAdd(&distinct, 1u);
            if (prev == kDensEmpty || prev == k[j]) break;
            h = (h + 1u) & static_cast<uint32_t>(kDensSlots - 1);
        }
        atomicAdd(&cnt[h], 1u);
    }
    __syncthreads();
    if (2u * distinct < runs) {
        for (int s = threadIdx.x; s < kDensSlots; s += 256) {
            const uint32_t c = cnt[s];
            if (c) __hip_atomic_fetch_add(D.counts + key[s], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kDensPer; ++j)
            if (k[j] != kDensEmpty) density_add_key(D.counts, k[j]);   // valid lanes: a prefix of the wave
    }
}

// volume = half(float(count) / 1e6) (Resources/Volume/volume.comp:14, stored
// to an r16f image), and the counters are cleared for the next count.  A
// block converts 2048 cells; each of its load and store instructions covers
// one contiguous span (a thread takes cells 4t..4t+3 and 1024+4t..1024+4t+3
// of the block), which measured 6.3 TB/s against 4.9 for 32 contiguous bytes
// per thread.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t density_half(uint32_t c)
{
    const _Float16 h = static_cast<_Float16>(__fdiv_rn(static_cast<float>(c), 1000000.0f));
    return __builtin_bit_cast(uint16_t, h);
}

__global__ __launch_bounds__(256) void k_density_volume(uint32_t* __restrict__ counts, uint16_t* __restrict__ vol,
                                                        size_t n2048)
{
    if (blockIdx.x >= n2048) return;
    u32x4* c = reinterpret_cast<u32x4*>(counts) + static_cast<size_t>(blockIdx.x) * 512;
    u32x2* v = reinterpret_cast<u32x2*>(vol) + static_cast<size_t>(blockIdx.x) * 512;
    const u32x4 a = __builtin_nontemporal_load(c + threadIdx.x);
    const u32x4 b = __builtin_nontemporal_load(c + 256 + threadIdx.x);
    const u32x2 oa = {density_half(a.x) | (density_half(a.y) << 16), density_half(a.z) | (density_half(a.w) << 16)};
    const u32x2 ob = {density_half(b.x) | (density_half(b.y) << 16), density_half(b.z) | (density_half(b.w) << 16)};
    __builtin_nontemporal_store(oa, v + threadIdx.x);
    __builtin_nontemporal_store(ob, v + 256 + threadIdx.x);
    const u32x4 z = {0u, 0u, 0u, 0u};
    c[threadIdx.x] = z;
    c[256 + threadIdx.x] = z;
}

// Several scenes, each through nframes frames of its window: particle
// records do not interact, so each thread carries its particle through the
// frames with u_t = u_hist + t * ldU and respawn frame P.frame + t, and
// stores it once.
__global__ __launch_bounds__(256) void k_trace_window_multi(const TraceScene* __restrict__ tab, int n, uint32_t per,
                                                            uint32_t blocks)
{
    const uint32_t blk = xcd_block(blockIdx.x, per);
    if (blk >= blocks) return;
    const TraceScene& ts = tab[scene_of_block(tab, n, blk)];
    const uint32_t id = (blk - ts.block0) * 256u + threadIdx.x;
    if (id >= ts.n) return;
    TraceParams P = ts.P;
    const gptr<f32x4> parts = as_global(reinterpret_cast<f32x4*>(ts.parts));
    const gptr<const float> uh = as_global(ts.u_hist);
    const f32x4 v = __builtin_nontemporal_load(&parts[id]);
    float4 p = make_float4(v.x, v.y, v.z, v.w);
    const uint32_t who = as_global(ts.pid)[id];   // loaded once (see k_trace_frames)
    const auto whofn = [who] { return who; };
    if (P.pow2) {
        for (uint32_t t = 0; t < ts.nframes; ++t) {
            p = trace_one<true>(P, uh + static_cast<size_t>(t) * ts.ldU, p, whofn, id, nullptr);
            P.frame++;
        }
    } else {
        for (uint32_t t = 0; t < ts.nframes; ++t) {
            p = trace_one<false>(P, uh + static_cast<size_t>(t) * ts.ldU, p, whofn, id, nullptr);
            P.frame++;
        }
    }
    const f32x4 o = {p.x, p.y, p.z, p.w};
    __builtin_nontemporal_store(o, &parts[id]);
}

__global__ void k_zero_int(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
hipError_t launch_contract_plain(int prec, const void* C, const double* omega, double* part, int K, hipStream_t s)
{
    const int S = plain_splits(K);
    const unsigned blocks = static_cast<unsigned>(K) * S;
    if (prec == PSF_C_FP64) {
        const size_t sh = sizeof(double) * K;
        if (K % 128 == 0)
            k_contract_plain<double, 2><<<blocks, 256, sh, s>>>(static_cast<const double*>(C), omega, part, K, S);
        else
            k_contract_plain<double, 1><<<blocks, 256, sh, s>>>(static_cast<const double*>(C), omega, part, K, S);
    } else {
        const size_t sh = sizeof(float) * K;
        if (K % 256 == 0)
            k_contract_plain<float, 4><<<blocks, 256, sh, s>>>(static_cast<const float*>(C), omega, part, K, S);
        else if (K % 128 == 0)
            k_contract_plain<float, 2><<<blocks, 256, sh, s>>>(static_cast<const float*>(C), omega, part, K, S);
        else
            k_contract_plain<float, 1><<<blocks, 256, sh, s>>>(static_cast<const float*>(C), omega, part, K, S);
    }
    return hipGetLastError();
}

hipError_t launch_update(double* omega, const double* part, const double* visc, const double* g, int K, double dt,
                         int gravity_active, int* status, double* hist, int hist_ld, int hist_t, hipStream_t s)
{
    const int S = plain_splits(K);
    if (K <= 256)
        k_update<1><<<1, 256, 0, s>>>(omega, part, S, visc, g, K, dt, gravity_active, status, hist, hist_ld, hist_t);
    else if (K <= 512)
        k_update<2><<<1, 256, 0, s>>>(omega, part, S, visc, g, K, dt, gravity_active, status, hist, hist_ld, hist_t);
    else if (K <= 1024)
        k_update<4><<<1, 256, 0, s>>>(omega, part, S, visc, g, K, dt, gravity_active, status, hist, hist_ld, hist_t);
    else
        k_update<16><<<1, 256, 0, s>>>(omega, part, S, visc, g, K, dt, gravity_active, status, hist, hist_ld, hist_t);
    return hipGetLastError();
}

// kernel<<<grid, block, 0, s>>>(args...), or with ev the same launch through
// hipExtLaunchKernelGGL, which records ev's events at the dispatch's start and end
#define PSF_LAUNCH(kernel, grid, block, s, ev, ...)                                                        \
    do {                                                                                                   \
        if (ev) hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, (ev)->start, (ev)->stop, 0,   \
                                      __VA_ARGS__);                                                        \
        else hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, __VA_ARGS__);                       \
    } while (0)

hipError_t launch_reconstruct(int prec, const void* Phi, size_t ldF, uint32_t F, int K, const double* omega,
                              float* u, hipStream_t s, const LaunchEvents* ev)
{
    if (prec == PSF_PHI_FP32_ACC32)   // one column of the FP32-accumulating window kernel (the same bits)
        return launch_reconstruct_window(prec, 1, 1, Phi, ldF, F, K, omega, u, s, ev);
    if (prec == PSF_PHI_FP32) {
        constexpr int V = kGemvVecF32;
        const unsigned grid = static_cast<unsigned>((ldF + 256 * V - 1) / (256 * V));
        PSF_LAUNCH((k_reconstruct<float, V>), grid, 256, s, ev, static_cast<const float*>(Phi), ldF, F, K, omega, u);
    } else {
        const unsigned grid = static_cast<unsigned>((ldF + 256 * 2 - 1) / (256 * 2));
        PSF_LAUNCH((k_reconstruct<double, 2>), grid, 256, s, ev, static_cast<const double*>(Phi), ldF, F, K, omega, u);
    }
    return hipGetLastError();
}

hipError_t launch_reconstruct_window(int prec, int hist_ld, int ncols, const void* Phi, size_t ldF, uint32_t F, int K,
                                     const double* hist, float* uh, hipStream_t s, const LaunchEvents* ev)
{
    if (ldF % 256 != 0) return hipErrorInvalidValue;
    if (ncols < 1 || ncols > 32 || hist_ld < ncols) return hipErrorInvalidValue;
    const unsigned grid = static_cast<unsigned>(ldF / 256);
    if (prec == PSF_PHI_FP32_ACC32) {
        const float* P = static_cast<const float*>(Phi);
        if (ncols > 16)
            PSF_LAUNCH((k_reconstruct_mfma<float, 2, float>), grid, 256, s, ev, P, ldF, F, K, hist, hist_ld, ncols, uh, ldF);
        else
            PSF_LAUNCH((k_reconstruct_mfma<float, 1, float>), grid, 256, s, ev, P, ldF, F, K, hist, hist_ld, ncols, uh, ldF);
    } else if (prec == PSF_PHI_FP32) {
        const float* P = static_cast<const float*>(Phi);
        if (ncols > 16)
            PSF_LAUNCH((k_reconstruct_mfma<float, 2>), grid, 256, s, ev, P, ldF, F, K, hist, hist_ld, ncols, uh, ldF);
        else
            PSF_LAUNCH((k_reconstruct_mfma<float, 1>), grid, 256, s, ev, P, ldF, F, K, hist, hist_ld, ncols, uh, ldF);
    } else {
        const double* P = static_cast<const double*>(Phi);
        if (ncols > 16)
            PSF_LAUNCH((k_reconstruct_mfma<double, 2>), grid, 256, s, ev, P, ldF, F, K, hist, hist_ld, ncols, uh, ldF);
        else
            PSF_LAUNCH((k_reconstruct_mfma<double, 1>), grid, 256, s, ev, P, ldF, F, K, hist, hist_ld, ncols, uh, ldF);
    }
    return hipGetLastError();
}

hipError_t launch_reconstruct_window_multi(int prec, int window, const WindowScene* d_tab, int n, uint32_t blocks,
                                           hipStream_t s)
{
    if (n <= 0 || blocks == 0) return hipSuccess;
    if (prec != PSF_PHI_FP32 && prec != PSF_PHI_FP32_ACC32) return hipErrorInvalidValue;
    if (window != 2 && window != 4 && window != 8 && window != 16 && window != 32) return hipErrorInvalidValue;
    if (prec == PSF_PHI_FP32_ACC32) {
        if (window == 32) k_reconstruct_mfma_pairs<2><<<blocks, 256, 0, s>>>(d_tab, n, window);
        else k_reconstruct_mfma_pairs<1><<<blocks, 256, 0, s>>>(d_tab, n, window);
    } else {
        if (window == 32) k_reconstruct_mfma_multi<2, double><<<blocks, 256, 0, s>>>(d_tab, n, window);
        else k_reconstruct_mfma_multi<1, double><<<blocks, 256, 0, s>>>(d_tab, n, window);
    }
    return hipGetLastError();
}

uint32_t window_blocks(int window, size_t ldF)
{
    (void)window;
    return static_cast<uint32_t>((ldF + 255) / 256);   // 256 faces per block
}

hipError_t launch_trace_window_multi(const TraceScene* d_tab, int n, uint32_t blocks, hipStream_t s)
{
    if (n <= 0 || blocks == 0) return hipSuccess;
    const uint32_t per = xcd_per(blocks);
    k_trace_window_multi<<<per * kXcds, 256, 0, s>>>(d_tab, n, per, blocks);
    return hipGetLastError();
}

hipError_t launch_gather_column(const double* hist, int ld, int t, int K, double* out, hipStream_t s)
{
    k_gather_column<<<(K + 255) / 256, 256, 0, s>>>(hist, ld, t, K, out);
    return hipGetLastError();
}

hipError_t launch_trace(const TraceParams& p, const float* u, float4* parts, const uint32_t* pid, uint32_t n,
                        int32_t* cells, const DensityGrid& D, hipStream_t s, const LaunchEvents* ev)
{
    if (n == 0) return hipSuccess;
    const uint32_t per = xcd_per((n + 255) / 256);
    if (p.pow2)
        PSF_LAUNCH(k_trace<true>, per * kXcds, 256, s, ev, p, u, parts, pid, n, cells, D, per);
    else
        PSF_LAUNCH(k_trace<false>, per * kXcds, 256, s, ev, p, u, parts, pid, n, cells, D, per);
    return hipGetLastError();
}

hipError_t launch_trace_frames(const TraceParams& p, const float* u0, size_t ldU, uint32_t nframes, float4* parts,
                               const uint32_t* pid, uint32_t n, const DensityGrid& D, hipStream_t s,
                               const LaunchEvents* ev)
{
    if (n == 0 || nframes == 0) return hipSuccess;
    const uint32_t per = xcd_per((n + 255) / 256);
    if (p.pow2)
        PSF_LAUNCH(k_trace_frames<true>, per * kXcds, 256, s, ev, p, u0, ldU, nframes, parts, pid, n, D, per);
    else
        PSF_LAUNCH(k_trace_frames<false>, per * kXcds, 256, s, ev, p, u0, ldU, nframes, parts, pid, n, D, per);
    return hipGetLastError();
}

hipError_t launch_density_count(const DensityGrid& D, const float4* parts, uint32_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_density_count<<<(n + kDensChunk - 1) / kDensChunk, 256, 0, s>>>(D, parts, n);
    return hipGetLastError();
}

hipError_t launch_density_volume(uint32_t* counts, uint16_t* volume, size_t n2048, hipStream_t s)
{
    if (n2048 == 0) return hipSuccess;
    k_density_volume<<<static_cast<unsigned>(n2048), 256, 0, s>>>(counts, volume, n2048);
    return hipGetLastError();
}

__global__ void k_fill_f64(double* p, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_f64(double* p, int n, double v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_fill_f64<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

// empty kernel that marks the ends of a timed region in kernel traces
__global__ void k_timed_region_mark(int) {}

hipError_t launch_timed_region_mark(int which, hipStream_t s)
{
    k_timed_region_mark<<<1, 64, 0, s>>>(which);
    return hipGetLastError();
}

hipError_t launch_zero_int(int* p, int n, hipStream_t s)
{
    k_zero_int<<<(n + 255) / 256, 256, 0, s>>>(p, n);
    return hipGetLastError();
}

}  // namespace psf
