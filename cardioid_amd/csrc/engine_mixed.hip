// engine_mixed.hip -- the mixed-precision fused step over the engine's slabs
// (gfx950).
//
// Same step as engine_f64.hip with the storage and arithmetic of
// tt06_mixed.hpp and an FP32 stencil on voltage differences,
//    dVm_D = scale * sum_{k=1..18} A_k * (phi_k - phi_0),
// which is the FGR operator itself (its centre weight is minus the sum of
// the others, elec/FGRDiffusionOMP.cc:214-219) without the centre weight.
// The weights A_k are held compactly: a float per tile and neighbour plus
// an FP16 difference per cell (slab.cpp::compact_weights).
#pragma clang fp contract(fast)

#include <hip/hip_fp16.h>

#include "engine.hpp"
#include "tt06_mixed.hpp"

namespace cdh {

using namespace mixed;

namespace {

// ref + the low (or high) FP16 half of a word, in one v_fma_mix_f32 with the
// inline constant 1.0: exact sum, rounded once, so ref + (-ref) is exactly 0
__device__ __forceinline__ float half_lo_plus(uint32_t packed, float ref)
{
   float r;
   asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(packed), "s"(ref));
   return r;
}
__device__ __forceinline__ float half_hi_plus(uint32_t packed, float ref)
{
   float r;
   asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
       : "=v"(r)
       : "v"(packed), "s"(ref));
   return r;
}

__device__ __forceinline__ void cell_mixed(const EngineArgs &a, const SlabDev &S, int sl,
                                           const SlabView &v, int blk, int lane)
{
   const int64_t i = (int64_t)blk * kBlk + lane;
   // padding cells of a slab's last tile compute on a copy of a tissue
   // cell's box point and store only their own states, so the tile runs
   // without a divergent exit
   const uint32_t meta = a.meta[i];
   const bool live = !(meta & kMetaPad);
   const uint32_t b = a.bi[i];
   const uint32_t b8 = b * 8u;
   const __amdgpu_buffer_rsrc_t R = box_rsrc(a.box_cur);
   // the cell's parameter set first: every later wait for it would also
   // wait for whatever was issued before it
   const ParamSet P = a.params[v.param_base + (int)(meta & 0xFFu)];
   __builtin_amdgcn_sched_barrier(0);
   const double V = box_load(R, b8, S.soff[0]);

   // states stream once per step (non-temporal), in the slab's runs of kRun
   // cells (run_index): the base addresses and the run width are wave-uniform
   // (buffer resources over the run's base, field s at SGPR offset
   // s * run width * field size, the lane's cell at VGPR offset)
   const int first = __builtin_amdgcn_readfirstlane(S.first_blk);
   const int64_t k0 = (int64_t)(blk - first) * kBlk;
   const int64_t r0 = k0 & ~(kRun - 1);
   const uint32_t rw =
       (uint32_t)run_width(k0, (int64_t)__builtin_amdgcn_readfirstlane(S.ntile) * kBlk);
   const int64_t c0 = (int64_t)first * kBlk + r0;
   const __amdgpu_buffer_rsrc_t R32 = stream_rsrc(a.st32 + c0 * N_STATE32);
   const __amdgpu_buffer_rsrc_t R64 = stream_rsrc(a.st64 + c0 * N_STATE64);
   const uint32_t lc = (uint32_t)(k0 - r0 + lane);
   float s32[N_STATE32];
#pragma unroll
   for (int s = 0; s < N_STATE32; s++) s32[s] = stream_load_f32(R32, 4u * lc, 4u * rw * s);
   double s64[N_STATE64];
#pragma unroll
   for (int s = 0; s < N_STATE64; s++) s64[s] = stream_load_f64(R64, 8u * lc, 8u * rw * s);
   __builtin_amdgcn_sched_barrier(0);
   // then the stencil's weights and neighbours: loads return in issue order,
   // so the reaction waits for the states only and the stencil's loads land
   // under its arithmetic; the sum comes after the reaction
   // the weight words in three loads (grp_index9): two of 16 and one of 4
   // bytes per lane (0.932 against 0.941 ms per step on the 20.1 M-cell
   // shell with one load per word, bit for bit the same)
   const uint32_t *wp = a.Aw16 + (int64_t)blk * (9 * kBlk);
   uint32_t hw[9];
   double nb[18];
#pragma unroll
   for (int g = 0; g < 2; g++) {
      const u32x4 q = __builtin_nontemporal_load((const u32x4 *)(wp + g * 4 * kBlk) + lane);
      hw[4 * g] = q.x;
      hw[4 * g + 1] = q.y;
      hw[4 * g + 2] = q.z;
      hw[4 * g + 3] = q.w;
   }
   hw[8] = __builtin_nontemporal_load(wp + 8 * kBlk + lane);
#pragma unroll
   for (int k = 0; k < 18; k++) nb[k] = box_load(R, b8, S.soff[k + 1]);
   __builtin_amdgcn_sched_barrier(0);

   const double K_i = P.K0 + (double)s32[M_DK_I];
   double iK;
   double dVR = 0.0;
   iK = 0.0;
   if (a.gate_mask & (1u << 12)) dVR = nongateM(*a.K, P, a.coef, a.dt, V, K_i, s32, s64, &iK);
   gatesM(a.coef, a.K->exp_coef, P.s_switch, (uint8_t)(meta >> 8), V, s32, a.gate_mask);

   // the stencil: weight (in units of 2^-wexp) = the tile's reference
   // (wave-uniform, scalar loads) plus the cell's FP16 difference, exact zero
   // for a zero weight; the unit goes into the diffusion scale
   const float *ref = a.Aref + (int64_t)blk * 18;
   float acc = 0.0f;
#pragma unroll
   for (int p = 0; p < 9; p++) {
      const float w0 = half_lo_plus(hw[p], ref[2 * p]);
      const float w1 = half_hi_plus(hw[p], ref[2 * p + 1]);
      acc = __builtin_fmaf(w0, (float)(nb[2 * p] - V), acc);
      acc = __builtin_fmaf(w1, (float)(nb[2 * p + 1] - V), acc);
   }
   const double dVD = (double)acc * (v.scale * (double)S.wdelta_scale);

   engine_range(a, S, v, sl, a.parity, live, lane, i, V, dVR, dVD);
   const double iStim = slab_stim(S, v, a.parity, meta);
   const double Vn = V + a.dt * (dVR + dVD + iStim);
   // K_i follows dVK_i = K_i/c9 + Vm of the reference: dK_i = c9*dt*(iK - dVm_D - iStim)
   s32[M_DK_I] = (float)((double)s32[M_DK_I] + (a.dt * a.K->c9) * (iK - dVD - iStim));
#pragma unroll
   for (int s = 0; s < N_STATE32; s++) stream_store_f32(s32[s], R32, 4u * lc, 4u * rw * s);
#pragma unroll
   for (int s = 0; s < N_STATE64; s++) stream_store_f64(s64[s], R64, 8u * lc, 8u * rw * s);
   uint32_t ev = 0;
   if (live) {
      a.box_next[b] = Vn;
      if (a.sensors) ev = engine_sensors(a, v, i, V, Vn, P.Vm0);
   }
   // every lane of the tile is here (padding lanes with no events)
   if (a.sensors & 1) sensor_count(a.counts + 4 * sl, ev);
}

}  // namespace

// Tile-to-XCD mapping as k_engine_f64.
template <int WAVES>
__global__ __launch_bounds__(256, WAVES) void k_engine_mixed(EngineArgs a)
{
   const int wpg = blockDim.x >> 6;
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int j = engine_tile_group(a) * wpg + wave;
   if (j >= a.ntile) return;
   const int blk = a.tiles ? (int)a.tiles[j] : a.tile_base + j;
   const int sl = a.blk_slab[blk];
   if (sl == kNoSlab) return;
   SlabDev &S = a.slabs[sl];
   if (!S.active) return;
   SlabView v;
   slab_view(S, a.parity, a.dt, v);
   cell_mixed(a, S, sl, v, blk, lane);
   if (blk == S.first_blk && lane == 0) slab_advance(S, a.parity, v);
}

template __global__ void k_engine_mixed<3>(EngineArgs);
template __global__ void k_engine_mixed<4>(EngineArgs);

}  // namespace cdh
