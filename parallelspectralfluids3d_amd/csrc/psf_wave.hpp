// Device helpers shared by the persistent reduced-step kernels: BF16
// rounding and fixed-order wave-wide FP64 sums.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psf {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return static_cast<uint16_t>(b >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Wave-wide FP64 sum on DPP lane moves (no LDS traffic): quad butterflies,
// half-row and row mirrors, then the row broadcasts of lanes 15 and 31; lane
// 63 ends with the total, which is broadcast.  The order is fixed, so every
// wave of every workgroup rounds identically.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_total(double v)
{
    v += dpp_move<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_move<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_move<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_move<0x140, 0xF>(v);   // row_mirror
    v += dpp_move<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}


}  // namespace psf
