// kernels.hpp -- kernel argument blocks and small device helpers shared by
// the FP64 and mixed-precision kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "box_access.hpp"
#include "tt06_model.hpp"

namespace cdh {

struct ParamSet;

constexpr int kMaxStim = 8;

// Reaction over n cells of one reaction group (drop-in object).
struct ReactionArgs {
   // the constants in device memory rather than a kernel argument: the
   // kernel would otherwise load all of them into SGPRs at entry and spill
   const TT06Const *K;
   const double *coef;          // 468 dt-scaled fit coefficients
   const ParamSet *params;      // parameter sets
   const uint8_t *pid;          // per-cell parameter set, or null (set 0)
   const uint8_t *flags;        // per-cell gate flags
   double *state;               // 19 states in 64-cell tiles (blk_index)
   int n;
   const int *index;            // reaction-local -> array index, or null (identity)
   const double *Vm;
   double *dVm;
   double dt;
   uint32_t gate_mask;          // bits 0..12 set; see gates64 and k_tt06_calc_pipe_f64
   const TT06ExactConst *E;     // configuration A (exact functions) constants, or null
   double *sink;                // 20 x 64 doubles: the pipelined kernel's stores of
                                // lanes past n (k_tt06_calc_pipe_f64)
};

struct StencilArgs {
   int n;
   const uint32_t *bi;          // box index per cell
   const double *box;           // Vm in the padded box
   const float *Aw;             // off-centre weights, NodeLocation order 1..18, in 64-cell
                                // tiles in load groups (grp_index18); the centre weight is
                                // minus their sum
   int off[19];                 // NodeLocation offsets in the box
   double scale;                // diffusionScale
   double *dVm;
   int groups_per_xcd;          // 256-cell blocks per XCD (grid = 8 * groups_per_xcd)
};

// Activation sensors, evaluated after the step at the new time.
//   tact0: first time Vm > 0            (ActivationTimeSensor.cc:99-111)
//   tact / trec: first up-crossing of the threshold, and the first
//   down-crossing after it                (ActivationAndRecoverySensor.cc:76-92)
struct SensorArgs {
   int n;
   int enabled;
   int first_eval;              // 1 on the first evaluation after the start
   double time;
   double threshold;
   double *tact0;
   double *tact;
   double *trec;
   uint8_t *active;             // threshold state (stateful form only)
   uint8_t *activated;          // Vm > 0 seen (stateful form only)
   // APD90 (no counterpart in the reference): after activation track the
   // largest Vm; t90 = first time Vm < Vrest + 0.1*(Vmax - Vrest), Vrest =
   // the cell type's rest Vm
   int apd;
   uint8_t *apd_phase;          // 0 waiting, 1 tracking Vmax, 2 done
   double *vmax;
   double *t90;
   uint32_t *counts;            // cells whose tact0, tact, trec have been recorded
                                // (3 counters; kept by the drop-in sensor kernels)
};

// Events of one sensor evaluation: which of the cell's times it recorded.
constexpr unsigned kEvTact0 = 1u, kEvTact = 2u, kEvTrec = 4u;

// Adds a wavefront's recorded times to counts[0..2] (tact0, tact, trec):
// one ballot per kind, one atomic per kind that occurred, from the lowest
// lane that took part.  Call from every lane that evaluated a cell (the
// others contribute nothing; lanes can hold several cells' events summed in
// `per_lane`, one count per kind, packed 8 bits each).
__device__ __forceinline__ void sensor_count(uint32_t *counts, uint32_t per_lane)
{
   const uint64_t any = __ballot(per_lane != 0);
   if (!any) return;
   const int lane = threadIdx.x & 63;
   for (int k = 0; k < 3; k++) {
      uint32_t c = (per_lane >> (8 * k)) & 0xFFu;
      for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
      if (c && lane == __builtin_ctzll(any)) atomicAdd(counts + k, c);
   }
}

// per-lane packing of one cell's events for sensor_count
__device__ __forceinline__ uint32_t sensor_events_packed(unsigned ev)
{
   return ((ev & kEvTact0) ? 1u : 0u) | ((ev & kEvTact) ? 1u << 8 : 0u) |
          ((ev & kEvTrec) ? 1u << 16 : 0u);
}

__device__ __forceinline__ void apd_eval(const SensorArgs &a, int i, double v, double vrest,
                                         bool activated_now)
{
   const uint8_t ph = a.apd_phase[i];
   if (ph == 0) {
      if (activated_now) {
         a.apd_phase[i] = 1;
         a.vmax[i] = v;
      }
   } else if (ph == 1) {
      const double vm = a.vmax[i];
      if (v > vm) {
         a.vmax[i] = v;
      } else if (v < vrest + 0.1 * (vm - vrest)) {
         a.t90[i] = a.time;
         a.apd_phase[i] = 2;
      }
   }
}

// Stateful form, the reference's own bookkeeping.
// The stateful form with the cell's two flags already loaded (activated,
// active); only a change of state writes.
// Returns the times it recorded (kEv* bits).
__device__ __forceinline__ unsigned sensor_eval_flags(const SensorArgs &a, int i, double v,
                                                      double vrest, uint8_t activated,
                                                      uint8_t active)
{
   unsigned ev = 0;
   if (!activated && v > 0) {
      a.activated[i] = 1;
      a.tact0[i] = a.time;
      ev |= kEvTact0;
   }
   if (active && v < a.threshold) {
      active = 0;
      a.active[i] = 0;
      if (a.trec[i] < 0 && a.tact[i] >= 0) {
         a.trec[i] = a.time;
         ev |= kEvTrec;
      }
   }
   bool now = false;
   if (!active && v > a.threshold) {
      a.active[i] = 1;
      if (a.tact[i] < 0) {
         a.tact[i] = a.time;
         now = true;
         ev |= kEvTact;
      }
   }
   if (a.apd) apd_eval(a, i, v, vrest, now);
   return ev;
}

__device__ __forceinline__ unsigned sensor_eval(const SensorArgs &a, int i, double v,
                                                double vrest)
{
   return sensor_eval_flags(a, i, v, vrest, a.activated[i], a.active[i]);
}

// Pair form from (Vm before, Vm after) of the step: the state before the
// step is "Vm above the level" (false before the first evaluation), which
// equals the stateful form unless Vm lands exactly on the level.
// Returns the times it recorded (kEv* bits).
__device__ __forceinline__ unsigned sensor_eval_pair(const SensorArgs &a, int i, double v0,
                                                     double v1, double vrest)
{
   unsigned ev = 0;
   const bool was0 = !a.first_eval && v0 > 0;
   if (!was0 && v1 > 0 && a.tact0[i] < 0) {
      a.tact0[i] = a.time;
      ev |= kEvTact0;
   }
   const bool wasT = !a.first_eval && v0 > a.threshold;
   bool now = false;
   if (!wasT && v1 > a.threshold && a.tact[i] < 0) {
      a.tact[i] = a.time;
      now = true;
      ev |= kEvTact;
   }
   if (wasT && v1 < a.threshold && a.trec[i] < 0 && a.tact[i] >= 0) {
      a.trec[i] = a.time;
      ev |= kEvTrec;
   }
   if (a.apd) apd_eval(a, i, v1, vrest, now);
   return ev;
}

// Voltage range check (Simulate::checkRanges, elec/Simulate.cc:35-53):
// cells whose Vm at the start of a step lies outside [vmin, vmax].  The
// reference prints one warning per such cell (Simulate::outOfRange, :55-74);
// on the device each wavefront adds its count with one atomic and the first
// wavefront to find one records it.
struct RangeRec {
   unsigned long long count;    // out-of-range (cell, step) pairs
   int32_t taken;               // 1 once the first record is written
   int32_t loop;                // loop counter of that step (before its increment)
   int64_t cell;                // engine cell index
   double V, dVmR, dVmD;
};

__device__ __forceinline__ void range_report(RangeRec *r, uint64_t mask, int lane, int loop,
                                             int64_t cell, double V, double dVmR, double dVmD)
{
   if (lane != __builtin_ctzll(mask)) return;
   atomicAdd(&r->count, (unsigned long long)__builtin_popcountll(mask));
   if (atomicCAS(&r->taken, 0, 1) == 0) {
      r->loop = loop;
      r->cell = cell;
      r->V = V;
      r->dVmR = dVmR;
      r->dVmD = dVmD;
   }
}

// kernels (kernels_f64.hip)
__global__ void k_tt06_calc_pipe_f64(ReactionArgs a);
__global__ void k_tt06_gate_f64(ReactionArgs a);
__global__ void k_tt06_calc_exact_f64(ReactionArgs a);
__global__ void k_tt06_gate_exact_f64(ReactionArgs a);
__global__ void k_scatter_f64(int n, const uint32_t *bi, const double *src, double *box);
__global__ void k_scatter4_f64(int n, const uint32_t *bi, const double *src, double *box);
__global__ void k_fgr_calc_f64(StencilArgs a);
__global__ void k_integrate_f64(int n, double dt, double *Vm, const double *dVmR,
                                const double *dVmD, const double *iStim);
__global__ void k_fill_f64(int n, double v, double *x);
__global__ void k_set_list_f64(int n, const int *list, double v, double *x);
__global__ void k_sensor_f64(SensorArgs a, const double *Vm, const double *vrest);
__global__ void k_sensor4_f64(SensorArgs a, const double *Vm, const double *vrest);
__global__ void k_check_range_f64(int n, const double *Vm, const double *dVmR, const double *dVmD,
                                  double vmin, double vmax, int loop, RangeRec *r);

}  // namespace cdh
