// box_access.hpp -- stencil neighbour loads through a buffer resource.
//
// The cell's box offset is the VGPR offset and the neighbour's offset
// (shifted non-negative by kBoxShift points) the SGPR offset of one buffer
// load, so none of the 19 neighbour addresses costs a VALU instruction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdh {

constexpr int64_t kBoxShift = int64_t(1) << 22;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t box_rsrc(const double *box)
{
   // raw buffer, no stride, the whole 32-bit range (the host keeps arena
   // boxes below 2^28 points); third dword as for every gfx9 target
   return __builtin_amdgcn_make_buffer_rsrc((void *)(box - kBoxShift), (short)0, -1, 0x00020000);
}

__device__ __forceinline__ double box_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
   return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

}  // namespace cdh
