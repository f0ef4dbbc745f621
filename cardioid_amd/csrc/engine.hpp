// engine.hpp -- the fused step over many tissue slabs (device side).
//
// The engine advances any number of independent slabs (anatomies with their
// own box, stimuli, drug scales and clock) in one launch per step.  Cells of
// all slabs live in one arena of 64-cell tiles; each tile belongs to one
// slab (blk_slab), so everything per slab is wave-uniform.  Nothing that
// changes from step to step is a kernel argument: a slab's time and loop
// count are double buffered in device memory by step parity, which is what
// lets a run of steps be captured once in a HIP graph and replayed.
//
// One step for one cell is the omp-loop step of elec/simulationLoop.cc:
// 170-241: stimulus (BoxStimulus.cc:53-64, PeriodicPulse.cc:22-28),
// reaction at Vm_t (TT06Dev_Reaction.cc:393-407), FGR stencil on Vm_t
// (FGRDiffusionOMP.cc:109-131), Vm += dt*(dVmR + dVmD + iStim)
// (simulationLoop.cc:221-227), time += dt (:230), then the activation
// sensors at the new time (ActivationTimeSensor.cc:99-111,
// ActivationAndRecoverySensor.cc:76-92).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "tt06_f64.hpp"
#include "tt06_model.hpp"

namespace cdh {

constexpr int kSlabStim = 8;                 // box stimuli per slab (mask bits)
constexpr uint16_t kNoSlab = 0xFFFF;         // tile not in use

// per-cell metadata word
constexpr uint32_t kMetaPad = 0x8000u;       // padding cell of a slab's last tile
__host__ __device__ inline uint32_t meta_pack(uint32_t pid, uint32_t flags, uint32_t stim)
{
   return (pid & 0xFFu) | ((flags & 0xFFu) << 8) | ((stim & 0xFFu) << 16);
}

// PeriodicPulse with its stimulus window (elec/PeriodicPulse.cc:22-28,
// elec/Stimulus.hh:21-27)
struct PulseDev {
   double t0, tf, period, phase, duration, amplitude;
};

// One slab.  A step of parity p reads time[p], loop[p] and the stimulus
// values of that time (stim_on[p], stim_val[p]); the slab's first tile
// writes time[p^1] = time[p] + dt, loop[p^1] = loop[p] + 1 and the stimulus
// values at time[p^1].
struct alignas(16) SlabDev {
   double time[2];
   int32_t loop[2];
   int32_t active;              // 0: the slab's tiles are skipped
   int32_t first_blk;           // the slab's first tile, which advances the clock
   int32_t ntile;               // the slab's tiles (mixed states: run_index over ntile * 64 cells)
   uint32_t soff[19];           // (offset_k + kBoxShift) * 8, NodeLocation order
   int32_t param_base;          // parameter sets of this slab start here
   int32_t nstim;
   uint32_t stim_on[2];         // bit k: stimulus k is nonzero at time[p]
   int32_t check_range;         // Vm range check on (Simulate::checkRanges)
   float wdelta_scale;          // mixed form: scale of the FP16 weight differences
   int32_t sensor_start;        // loop of the sensors' first evaluation's step: they have
                                // seen no Vm before it (fresh sensors, at the start or a
                                // restart); -1 while they continue a run
   int32_t pad0_;
   double scale;                // diffusionScale
   double threshold;            // activation threshold
   double vm_min, vm_max;       // the check's range
   double stim_val[2][kSlabStim];
   PulseDev stim[kSlabStim];
};

struct EngineArgs {
   // constants in device memory rather than kernel arguments: the kernel
   // would otherwise load all of them into SGPRs at entry and spill them
   const TT06Const *K;
   const TT06ExactConst *E;     // configuration A's constants (exact engines), or null
   const double *coef;
   const ParamSet *params;
   const uint32_t *meta;        // per cell: meta_pack(pid, flags, stim mask)
   const uint32_t *bi;          // per cell: index in the arena's box
   const uint16_t *blk_slab;    // per 64-cell tile: slab, or kNoSlab
   SlabDev *slabs;
   double *state;               // FP64: 19 fields in 64-cell blocks
   float *st32;                 // mixed: 17 fields, per slab in runs
   double *st64;                // mixed: Ca_i, Na_i, per slab in runs
   const float *Aw;             // FP64: 18 off-centre weights in 64-cell blocks
   const float *Aref;           // mixed: per tile, 18 reference weights
   const uint32_t *Aw16;        // mixed: per cell, FP16 differences, 9 words in 64-cell blocks
   const double *box_cur;       // Vm in the slabs' padded boxes (read)
   double *box_next;            // (written)
   int parity;
   const uint32_t *tiles;       // the live slabs' 64-cell tiles, in arena order; null when
                                // they are the one range tile_base .. tile_base + ntile - 1
   int tile_base;
   int ntile;                   // live tiles
   int groups_per_xcd;          // workgroups per XCD (grid = 8 * groups_per_xcd)
   int xcd_chunk_log2;          // tile order: -1 contiguous per XCD, else runs of 2^this
                                // workgroups dealt to the XCDs in turn
   double dt;
   uint32_t gate_mask;          // bits 0..12 set; see gates64 and the kernels
   int sensors;                 // bit 0: activation times, bit 1: APD90
   double *tact0, *tact, *trec;
   uint8_t *apd_phase;
   double *vmax, *t90;
   uint32_t *counts;            // per slab, 4 words: cells whose tact0, tact, trec have
                                // been recorded (a wavefront ballot and one atomic each)
   RangeRec *ranges;            // per slab
};

// PeriodicPulse::eval at `time` inside the window t0 < time < tf, else 0
// (host and device round alike: no contraction, IEEE division).
__host__ __device__ inline double pulse_value(const PulseDev &b, double time)
{
#pragma clang fp contract(off)
   if (!(time > b.t0 && time < b.tf)) return 0.0;
   const double t = time - floor(time / b.period) * b.period;
   if (t >= b.phase && t < b.phase + b.duration) return b.amplitude;
   return 0.0;
}

// Stimulus values of every stimulus of a slab at `time`.
__host__ __device__ inline void slab_stim_eval(const PulseDev *stim, int nstim, double time,
                                               uint32_t *on, double *val)
{
   *on = 0;
   for (int k = 0; k < kSlabStim; k++) {
      val[k] = 0.0;
      if (k < nstim) {
         val[k] = pulse_value(stim[k], time);
         if (val[k] != 0.0) *on |= 1u << k;
      }
   }
}

// Wave-uniform values of the slab a tile belongs to.
struct SlabView {
   double time1;                // time after the step
   int first_eval;
   uint32_t stim_on;
   double scale, threshold;
   int param_base;
   int check_range;             // the range check's switch and limits, read here at the
   double vm_min, vm_max;       // kernel's start: read after the state stores, they
                                // became vector loads that waited for those stores
};

__device__ __forceinline__ void slab_view(const SlabDev &S, int parity, double dt, SlabView &v)
{
#pragma clang fp contract(off)
   v.time1 = S.time[parity] + dt;
   v.first_eval = S.loop[parity] == S.sensor_start;
   v.scale = S.scale;
   v.threshold = S.threshold;
   v.param_base = S.param_base;
   v.stim_on = S.stim_on[parity];
   v.check_range = S.check_range;
   v.vm_min = S.vm_min;
   v.vm_max = S.vm_max;
}

// The cell's stimulus; later stimuli overwrite earlier ones, as BoxStimulus
// assigns (elec/BoxStimulus.cc:61-62), so the highest set bit wins.
__device__ __forceinline__ double slab_stim(const SlabDev &S, const SlabView &v, int parity,
                                            uint32_t meta)
{
   const uint32_t on = (meta >> 16) & v.stim_on & 0xFFu;
   if (!on) return 0.0;
   return S.stim_val[parity][31 - __builtin_clz(on)];
}

// The first tile's lane 0 advances the slab's clock after the step.
__device__ __forceinline__ void slab_advance(SlabDev &S, int parity, const SlabView &v)
{
   const int q = parity ^ 1;
   S.time[q] = v.time1;
   S.loop[q] = S.loop[parity] + 1;
   uint32_t on;
   double val[kSlabStim];
   slab_stim_eval(S.stim, S.nstim, v.time1, &on, val);
   for (int k = 0; k < kSlabStim; k++) S.stim_val[q][k] = val[k];
   S.stim_on[q] = on;
}

// Sensors at the time after the step, pair form (Vm before and after);
// returns the cell's recorded times packed for sensor_count.
__device__ __forceinline__ uint32_t engine_sensors(const EngineArgs &a, const SlabView &v,
                                                   int64_t i, double v0, double v1, double vrest)
{
   SensorArgs s;
   s.n = 0;
   s.enabled = a.sensors;
   s.first_eval = v.first_eval;
   s.time = v.time1;
   s.threshold = v.threshold;
   s.tact0 = a.tact0;
   s.tact = a.tact;
   s.trec = a.trec;
   s.active = nullptr;
   s.activated = nullptr;
   s.apd = (a.sensors >> 1) & 1;
   s.apd_phase = a.apd_phase;
   s.vmax = a.vmax;
   s.t90 = a.t90;
   s.counts = nullptr;
   return sensor_events_packed(sensor_eval_pair(s, (int)i, v0, v1, vrest));
}

// Vm range check of a wavefront (cells whose Vm at the step's start is out
// of range); uniform branch, one atomic per wavefront that finds any.
__device__ __forceinline__ void engine_range(const EngineArgs &a, const SlabDev &S,
                                             const SlabView &v, int sl, int parity, bool live,
                                             int lane, int64_t i, double V, double dVR,
                                             double dVD)
{
   if (!v.check_range) return;
   const bool out = live && (V > v.vm_max || V < v.vm_min);
   const uint64_t m = __ballot(out);
   if (m) range_report(a.ranges + sl, m, lane, S.loop[parity], i, V, dVR, dVD);
}

// The tile group (workgroup-sized run of tiles) of this workgroup.
// Workgroups are dealt round robin to the 8 XCDs (workgroup w on XCD w mod
// 8), and each XCD has its own L2.  Contiguous order (xcd_chunk_log2 < 0):
// XCD x serves the x-th eighth of the tile groups, so that its L2 holds the
// box rows its cells read.  Chunked order: runs of G = 2^xcd_chunk_log2
// consecutive tile groups go to the XCDs in turn.  Both cover every group
// once; the grid may hold a few workgroups past the last group, which exit.
__device__ __forceinline__ int engine_tile_group(const EngineArgs &a)
{
   const int xcd = blockIdx.x & 7, r = blockIdx.x >> 3;
   if (a.xcd_chunk_log2 < 0) return xcd * a.groups_per_xcd + r;
   const int sh = a.xcd_chunk_log2;
   const int c = r >> sh;
   return (((c << 3) + xcd) << sh) + (r - (c << sh));
}

// the FP64 engine's reaction arithmetic (engine_f64.hip)
enum EngineKind : int { kFits = 0, kFitsFma = 1, kExact = 2 };

// kernels (engine_f64.hip, engine_mixed.hip); WAVES = target waves per SIMD
template <int WAVES, int KIND>
__global__ void k_engine_f64(EngineArgs a);
template <int WAVES>
__global__ void k_engine_mixed(EngineArgs a);

}  // namespace cdh
