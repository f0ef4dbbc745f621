// engine_f64.hip -- the fused FP64 step over the engine's slabs (gfx950).
//
// Compiled with -ffp-contract=off: with FMA = false the reaction, the
// stencil and the Vm update round as the reference CPU path does, operation
// by operation (TT06Dev_Reaction::calc, elec/TT06Dev_Reaction.cc:393-407;
// FGRDiffusionOMP::calc, elec/FGRDiffusionOMP.cc:109-131; the integrator,
// elec/simulationLoop.cc:221-227).  KIND selects the reaction's arithmetic:
// kFits, the 33 rational fits as the reference rounds them; kFitsFma, the
// fit polynomials with fused multiply-adds (half the instructions of the
// multiply-add pairs) and everything else as it is; kExact, configuration A
// (TT06Dev tolerance 0, the factory's default deck): the exact voltage
// functions with the device's exp, log, sqrt and pow and the generic gate
// update (elec/TT06Func.cc:64-161), rounded as the drop-in reaction object
// rounds them, so that the fused step equals the unfused loop bit for bit.
#pragma clang fp contract(off)

#include "engine.hpp"
#include "tt06_exact.hpp"

namespace cdh {

namespace {

template <int KIND>
__device__ __forceinline__ void cell_f64(const EngineArgs &a, const SlabDev &S, int sl,
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
   const double V = box_load(R, b8, S.soff[0]);

   // the stencil first (FGRDiffusionOMP::calc order: A0*phi0, then
   // A_k*phi_k for k = 1..18): its neighbour and weight loads are issued with
   // the state loads at the start, so the reaction's arithmetic covers them.
   // The centre weight is minus the double sum of the 18 float weights in
   // order (FGRDiffusionOMP.cc:214-219), recomputed here bit for bit rather
   // than streamed from memory.
   // (one load per weight: the drop-in stencil's grouped loads, grp_index18,
   // measured no faster here, 1.516 against 1.513 ms per FMA step)
   const float *wp = a.Aw + (int64_t)blk * (18 * kBlk) + lane;
   float w[18];
   double A0 = 0.0;
#pragma unroll
   for (int k = 0; k < 18; k++) {
      w[k] = __builtin_nontemporal_load(wp + kBlk * k);
      A0 -= w[k];
   }
   double tmp = A0 * V;
#pragma unroll
   for (int k = 1; k < 19; k++) tmp += (double)w[k - 1] * box_load(R, b8, S.soff[k]);
   const double dVD = tmp * v.scale;

   // states stream once per step (non-temporal), in the slab's runs of kRun
   // cells with the fields in pairs (run_pair_index): a buffer over the
   // run's base, the pair's offset in an SGPR, the lane's cell in a VGPR
   const int first = __builtin_amdgcn_readfirstlane(S.first_blk);
   const int64_t k0 = (int64_t)(blk - first) * kBlk;
   const int64_t r0 = k0 & ~(kRun - 1);
   const uint32_t rw =
       (uint32_t)run_width(k0, (int64_t)__builtin_amdgcn_readfirstlane(S.ntile) * kBlk);
   const __amdgpu_buffer_rsrc_t RS = stream_rsrc(a.state + ((int64_t)first * kBlk + r0) * N_STATE);
   const uint32_t lc = (uint32_t)(k0 - r0 + lane);
   static_assert(N_STATE % 2 == 1, "the last state field stands alone after the pairs");
   double st[N_STATE];
#pragma unroll
   for (int s = 0; s + 1 < N_STATE; s += 2)
      stream_load_f64x2(RS, 16u * lc, 8u * rw * s, st[s], st[s + 1]);
   st[N_STATE - 1] = stream_load_f64(RS, 8u * lc, 8u * rw * (N_STATE - 1));
   const ParamSet P = a.params[v.param_base + (int)(meta & 0xFFu)];
   // the non-gate update in a basic block of its own (gate_mask bit 12,
   // always set): otherwise the compiler hoists its scalar coefficient loads
   // to the kernel entry and spills them
   double dVR = 0.0;
   if constexpr (KIND == kExact) {
      if (a.gate_mask & (1u << 12))
         dVR = nongate_core64(*a.K, P, ExactVoltageFns{a.E, a.K}, a.dt, V, st);
      if (!((meta >> 8) & kFlagGatesFixed))
         gates_exact64(*a.E, P.s_switch, a.dt, V, st, a.gate_mask);
   } else {
      constexpr bool FMA = KIND == kFitsFma;
      if (a.gate_mask & (1u << 12)) dVR = nongate64<FMA>(*a.K, P, a.coef, a.dt, V, st);
      gates64<FMA>(a.coef, a.K->exp_coef, P.s_switch, (uint8_t)(meta >> 8), V, st, a.gate_mask);
   }
   // Reads of the slab's descriptor after a vector store become vector loads
   // that wait for it, so the range check's limits come from the view read
   // at the start.  Since then the stores' place makes no difference: on one
   // MI355X, both instances within 0.2 % with the stores before or after the
   // range check and the stimulus (the FMA step 0.5 % faster than when the
   // limits were read here).
   if constexpr (KIND == kFitsFma) {
#pragma unroll
      for (int s = 0; s + 1 < N_STATE; s += 2)
         stream_store_f64x2(st[s], st[s + 1], RS, 16u * lc, 8u * rw * s);
      stream_store_f64(st[N_STATE - 1], RS, 8u * lc, 8u * rw * (N_STATE - 1));
   }
   engine_range(a, S, v, sl, a.parity, live, lane, i, V, dVR, dVD);
   const double iStim = slab_stim(S, v, a.parity, meta);
   const double Vn = V + a.dt * (dVR + dVD + iStim);
   if constexpr (KIND != kFitsFma) {
#pragma unroll
      for (int s = 0; s + 1 < N_STATE; s += 2)
         stream_store_f64x2(st[s], st[s + 1], RS, 16u * lc, 8u * rw * s);
      stream_store_f64(st[N_STATE - 1], RS, 8u * lc, 8u * rw * (N_STATE - 1));
   }
   uint32_t ev = 0;
   if (live) {
      a.box_next[b] = Vn;
      if (a.sensors) ev = engine_sensors(a, v, i, V, Vn, P.Vm0);
   }
   // every lane of the tile is here (padding lanes with no events)
   if (a.sensors & 1) sensor_count(a.counts + 4 * sl, ev);
}

}  // namespace

// Workgroups map to tiles in the XCD-aware order of engine_tile_group.  The
// grid covers the live slabs' tiles only: tiles of removed slabs cost nothing.
// When the live tiles are one range (one anatomy, a batch without holes)
// the tile is the index plus a base, with no list to read first.
template <int WAVES, int KIND>
__global__ __launch_bounds__(256, WAVES) void k_engine_f64(EngineArgs a)
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
   cell_f64<KIND>(a, S, sl, v, blk, lane);
   if (blk == S.first_blk && lane == 0) slab_advance(S, a.parity, v);
}

template __global__ void k_engine_f64<2, kFits>(EngineArgs);
template __global__ void k_engine_f64<3, kFits>(EngineArgs);
template __global__ void k_engine_f64<2, kFitsFma>(EngineArgs);
template __global__ void k_engine_f64<3, kFitsFma>(EngineArgs);
template __global__ void k_engine_f64<2, kExact>(EngineArgs);
template __global__ void k_engine_f64<3, kExact>(EngineArgs);

}  // namespace cdh
