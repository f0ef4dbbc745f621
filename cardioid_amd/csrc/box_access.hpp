// box_access.hpp -- stencil neighbour loads and state streams through
// buffer resources.
//
// The cell's box offset is the VGPR offset and the neighbour's offset
// (shifted non-negative by kBoxShift points) the SGPR offset of one buffer
// load, so none of the 19 neighbour addresses costs a VALU instruction.
// The mixed engine's state streams work alike: the run's base address sits
// in the resource, the lane's offset in a VGPR and the field's offset (its
// stride is the run width, not a constant) in an SGPR, so no field address
// is held in VGPRs across the reaction.
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

// a raw buffer over a wave-uniform base address (offsets below 2^32 bytes)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t stream_rsrc(const void *base)
{
   return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, -1, 0x00020000);
}

// non-temporal (cache policy NT) loads and stores of one field
constexpr int kNT = 2;
__device__ __forceinline__ double stream_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                  uint32_t soff)
{
   return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kNT));
}
__device__ __forceinline__ void stream_store_f64(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 uint32_t soff)
{
   typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));
   __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, kNT);
}
// two adjacent FP64 fields of a cell (a pair of run_pair_index), 16 bytes
__device__ __forceinline__ void stream_load_f64x2(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                  uint32_t soff, double &a, double &b)
{
   typedef unsigned int u32x4 __attribute__((__vector_size__(4 * sizeof(unsigned int))));
   const u32x4 v =
       __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kNT));
   a = __builtin_bit_cast(double, ((uint64_t)v[1] << 32) | v[0]);
   b = __builtin_bit_cast(double, ((uint64_t)v[3] << 32) | v[2]);
}
__device__ __forceinline__ void stream_store_f64x2(double a, double b, __amdgpu_buffer_rsrc_t r,
                                                   uint32_t voff, uint32_t soff)
{
   typedef unsigned int u32x4 __attribute__((__vector_size__(4 * sizeof(unsigned int))));
   const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
   u32x4 v;
   v[0] = (uint32_t)ua;
   v[1] = (uint32_t)(ua >> 32);
   v[2] = (uint32_t)ub;
   v[3] = (uint32_t)(ub >> 32);
   __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, kNT);
}
__device__ __forceinline__ float stream_load_f32(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 uint32_t soff)
{
   return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kNT));
}
__device__ __forceinline__ void stream_store_f32(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 uint32_t soff)
{
   __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, kNT);
}

}  // namespace cdh
