// kernels_f64.hip -- FP64 kernels of the drop-in path (gfx950).
//
// Compiled with -ffp-contract=off: every kernel here reproduces the
// reference CPU path's rounding, operation by operation.
//
//   k_tt06_calc_pipe_f64  TT06Dev_Reaction::calc        elec/TT06Dev_Reaction.cc:393-407
//   k_tt06_calc_exact_f64  the same in configuration A (exact functions,
//                       generic gates)                  elec/TT06Func.cc:64-161
//   k_scatter_f64       FGRDiffusionOMP::updateLocalVoltage  elec/FGRDiffusionOMP.cc:75-89
//   k_fgr_calc_f64      FGRDiffusionOMP::calc           elec/FGRDiffusionOMP.cc:109-131
//   k_integrate_f64     Vm += dt*(dVmR+dVmD+iStim)      elec/simulationLoop.cc:221-227
//   k_fill_f64 / k_set_list_f64   iStim = 0; BoxStimulus  elec/simulationLoop.cc:186-193,
//                                                   elec/BoxStimulus.cc:53-64
//   k_sensor_f64        ActivationTimeSensor / ActivationAndRecoverySensor eval
#pragma clang fp contract(off)

#include "kernels.hpp"
#include "row_shift.hpp"
#include "tt06_exact.hpp"
#include "tt06_f64.hpp"

namespace cdh {

// The drop-in objects keep their states and weights in 64-cell tiles:
// field q of cell i at blk_index(q, i, nq), a wavefront's fields one
// contiguous run.
//
// TT06Dev_Reaction::calc, software pipelined.  The grid holds only as many
// wavefronts as are resident at once (G); wavefront w walks the tiles
// w, w + G, w + 2G, ... and issues the loads of its next tile (its 19
// states, gate flags and Vm, and the index of the tile after) before the
// arithmetic of the current one, so that its memory traffic runs under its
// own arithmetic rather than only under other wavefronts'.  The loads are
// unconditional (past the last tile the current one is read again and the
// values dropped), which keeps the compiler's wait for the current tile's
// data from also waiting on the next tile's.  The stores are unconditional
// too: lanes past n (the last tile's padding; their states are padded to a
// whole tile, and they read Vm[0]) store into the object's 20 x 64-double
// sink instead of its states and dVm.  A store block skipped under a lane
// mask would make the compiler's waits assume no stores were issued and
// wait for every store before touching the next tile.  One parameter set
// per object (pid is not used).  Against one tile per wavefront with
// every load up front (4 waves/SIMD at 108 VGPRs), this form (3 waves/SIMD)
// took the reference-structured step on the 20.1 M-cell shell from 2.28 to
// 2.10 ms on one MI355X, bit for bit the same states, Vm and activation.
//
// The body takes its arrays as restrict-qualified arguments: the stores of
// one tile then provably leave the constants and fit coefficients alone,
// and the next tile's reads of those stay scalar loads instead of vector
// loads that would wait on every outstanding access.
namespace {

__device__ __forceinline__ void calc_pipe(const TT06Const *__restrict__ K,
                                          const double *__restrict__ coef,
                                          const ParamSet *__restrict__ params,
                                          const uint8_t *__restrict__ flags,
                                          double *__restrict__ state, const int *__restrict__ index,
                                          const double *__restrict__ Vm, double *__restrict__ dVm,
                                          double *__restrict__ sink, int n, double dt,
                                          uint32_t gate_mask)
{
   const int lane = threadIdx.x & 63;
   const int G = __builtin_amdgcn_readfirstlane(gridDim.x * (blockDim.x >> 6));
   const int ntile = (n + kBlk - 1) / kBlk;
   int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
   if (t >= ntile) return;
   const ParamSet P = params[0];
   auto cell_index = [&](int tt) -> int {
      const int i = tt * kBlk + lane;
      return i < n ? (index ? index[i] : i) : 0;
   };
   auto load_states = [&](int tt, double *st) {
      const double *sp = state + blk_base((int64_t)tt * kBlk + lane, N_STATE);
#pragma unroll
      for (int s = 0; s < N_STATE; s++) st[s] = __builtin_nontemporal_load(sp + kBlk * s);
   };
   // this tile's data, and the next tile's index
   int e = cell_index(t);
   int tn = t + G < ntile ? t + G : t;
   int en = cell_index(tn);
   double V = Vm[e];
   uint8_t fl = flags[t * kBlk + lane];
   double sa[N_STATE], sb[N_STATE];
   load_states(t, sa);
   // One tile: the next tile's loads into nxt (and the index of the one
   // after), then the arithmetic and stores of the current tile, held in
   // cur.  Returns false after the last tile.  The loop alternates the two
   // state buffers rather than copying one into the other.
   auto tile = [&](double *cur, double *nxt) -> bool {
      const int tnn = tn + G < ntile ? tn + G : tn;
      const int enn = cell_index(tnn);
      const double Vn = Vm[en];
      const uint8_t fln = flags[tn * kBlk + lane];
      load_states(tn, nxt);
      double dVR = 0.0;
      if (gate_mask & (1u << 12)) dVR = nongate64(*K, P, coef, dt, V, cur);
      gates64(coef, K->exp_coef, P.s_switch, fl, V, cur, gate_mask);
      const int i = t * kBlk + lane;
      const bool live = i < n;
      double *sp = live ? state + blk_base(i, N_STATE) : sink + lane;
#pragma unroll
      for (int s = 0; s < N_STATE; s++) __builtin_nontemporal_store(cur[s], sp + kBlk * s);
      *(live ? dVm + e : sink + kBlk * N_STATE + lane) = dVR;
      if (tn == t) return false;
      t = tn;
      tn = tnn;
      e = en;
      en = enn;
      V = Vn;
      fl = fln;
      return true;
   };
   while (tile(sa, sb) && tile(sb, sa)) {
   }
}

}  // namespace

__global__ __launch_bounds__(256, 3) void k_tt06_calc_pipe_f64(ReactionArgs a)
{
   calc_pipe(a.K, a.coef, a.params, a.flags, a.state, a.index, a.Vm, a.dVm, a.sink, a.n, a.dt,
             a.gate_mask);
}

// Gate update only (Reaction::updateGate of the pdr loop): the 12 gates at Vm.
__global__ __launch_bounds__(256) void k_tt06_gate_f64(ReactionArgs a)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= a.n) return;
   const int e = a.index ? a.index[i] : i;
   const double V = a.Vm[e];
   double st[N_STATE];
#pragma unroll
   for (int s = S_M; s < N_STATE; s++) st[s] = a.state[blk_index(s, i, N_STATE)];
   const ParamSet P = a.params[a.pid ? a.pid[i] : 0];
   gates64(a.coef, a.K->exp_coef, P.s_switch, a.flags[i], V, st, a.gate_mask);
#pragma unroll
   for (int s = S_M; s < N_STATE; s++) a.state[blk_index(s, i, N_STATE)] = st[s];
}

// Configuration A: the non-gate update with the exact fv0..fv6, then the
// generic gate update with the exact steady states and rates, at the
// step's Vm.  Gate flags bit 0 holds every gate of a cell (the pdr loop's
// updateNonGate); the generic loop has no v2 loop bounds, so otherwise every
// cell's gates advance.
__global__ __launch_bounds__(256) void k_tt06_calc_exact_f64(ReactionArgs a)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= a.n) return;
   const int e = a.index ? a.index[i] : i;
   const double V = a.Vm[e];
   double st[N_STATE];
#pragma unroll
   for (int s = 0; s < N_STATE; s++) st[s] = a.state[blk_index(s, i, N_STATE)];
   const ParamSet P = a.params[0];
   double dVR = 0.0;
   if (a.gate_mask & (1u << 12)) dVR = nongate_core64(*a.K, P, ExactVoltageFns{a.E, a.K}, a.dt, V, st);
   if (!(a.flags[i] & kFlagGatesFixed)) gates_exact64(*a.E, P.s_switch, a.dt, V, st, a.gate_mask);
#pragma unroll
   for (int s = 0; s < N_STATE; s++) a.state[blk_index(s, i, N_STATE)] = st[s];
   if (a.dVm) a.dVm[e] = dVR;
}

// Configuration A gate update only (Reaction::updateGate of the pdr loop).
__global__ __launch_bounds__(256) void k_tt06_gate_exact_f64(ReactionArgs a)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= a.n) return;
   const int e = a.index ? a.index[i] : i;
   const double V = a.Vm[e];
   double st[N_STATE];
#pragma unroll
   for (int s = S_M; s < N_STATE; s++) st[s] = a.state[blk_index(s, i, N_STATE)];
   gates_exact64(*a.E, a.params[0].s_switch, a.dt, V, st, a.gate_mask);
#pragma unroll
   for (int s = S_M; s < N_STATE; s++) a.state[blk_index(s, i, N_STATE)] = st[s];
}

__global__ __launch_bounds__(256) void k_scatter_f64(int n, const uint32_t *__restrict__ bi,
                                                     const double *__restrict__ src,
                                                     double *__restrict__ box)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) box[bi[i]] = src[i];
}

// The same, four cells per thread: one 16-byte load of box indices and two
// of values per thread (bi and src 16-byte aligned; the host checks).
__global__ __launch_bounds__(256) void k_scatter4_f64(int n, const uint32_t *__restrict__ bi,
                                                      const double *__restrict__ src,
                                                      double *__restrict__ box)
{
   const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
   if (i + 3 < n) {
      const uint4 b = *reinterpret_cast<const uint4 *>(bi + i);
      const double2 s0 = *reinterpret_cast<const double2 *>(src + i);
      const double2 s1 = *reinterpret_cast<const double2 *>(src + i + 2);
      box[b.x] = s0.x;
      box[b.y] = s0.y;
      box[b.z] = s1.x;
      box[b.w] = s1.y;
   } else {
      for (int64_t k = i; k < n; k++) box[bi[k]] = src[k];
   }
}

// the drop-in box's neighbour loads, for gather_rows
struct DropInBox {
   const double *box;
   const int *off;
   __device__ __forceinline__ double at(uint32_t b, int k) const { return box[(int64_t)b + off[k]]; }
   __device__ __forceinline__ uint32_t rel(int k) const { return (uint32_t)off[k]; }
   __device__ __forceinline__ double at_rel(uint32_t b, uint32_t r) const
   {
      return box[(int64_t)b + (int32_t)r];
   }
};

// Blocks are dealt round-robin to the 8 XCDs; each XCD takes a contiguous
// range of cells, so that its L2 holds the box rows and planes they read.
__global__ __launch_bounds__(256) void k_fgr_calc_f64(StencilArgs a)
{
   const int xcd = blockIdx.x & 7;
   const int64_t i =
       ((int64_t)xcd * a.groups_per_xcd + (blockIdx.x >> 3)) * blockDim.x + threadIdx.x;
   // a wavefront past the last cell has nothing to do; in the last one,
   // lanes past it work on the last cell and store nothing, so that every
   // lane takes part in the lane exchanges of the neighbour loads
   if ((i & ~(int64_t)(kBlk - 1)) >= a.n) return;
   const bool act = i < a.n;
   const int64_t ic = act ? i : a.n - 1;
   // Every load is issued before any arithmetic: the box index, then the
   // weights (they stream while the index arrives), then the neighbours.
   // Left to itself the compiler interleaves the neighbour loads with waits
   // on the weights for the centre-weight sum, which costs a third round
   // trip to memory per wave (4 % of the kernel's time).
   const uint32_t b = a.bi[ic];
   // the 18 weights in five loads (grp_index18): four of 16 bytes and one
   // of 8 bytes per lane.  The kernel's texture units are busy for 95 % of
   // its time (rocprof TA_TA_BUSY, TD_TD_BUSY), so fewer load instructions
   // help: 365 against 373 us per call on the 20.1 M-cell shell with one
   // load per weight, bit for bit the same
   const float *tw = a.Aw + (ic >> 6) * (18 * kBlk);
   const int lane = (int)(i & (kBlk - 1));
   float w[18];
#pragma unroll
   for (int g = 0; g < 4; g++) {
      const f32x4 q = __builtin_nontemporal_load((const f32x4 *)(tw + g * 4 * kBlk) + lane);
      w[4 * g] = q.x;
      w[4 * g + 1] = q.y;
      w[4 * g + 2] = q.z;
      w[4 * g + 3] = q.w;
   }
   {
      const f32x2 q = __builtin_nontemporal_load((const f32x2 *)(tw + 16 * kBlk) + lane);
      w[16] = q.x;
      w[17] = q.y;
   }
   __builtin_amdgcn_sched_barrier(0);
   // the 19 neighbours: nine loads, the ten x-1 / x+1 values from the
   // neighbouring lanes where they hold them (row_shift.hpp)
   double v[19];
   gather_rows<kRowsX>(DropInBox{a.box, a.off}, b, act, lane, v);
   __builtin_amdgcn_sched_barrier(0);
   // the weights stream once per step; the centre weight is minus their sum
   // in order, in double (FGRDiffusionOMP.cc:214-219), recomputed bit for bit
   double A0 = 0.0;
#pragma unroll
   for (int k = 0; k < 18; k++) A0 -= w[k];
   double tmp = A0 * v[0];
#pragma unroll
   for (int k = 1; k < 19; k++) tmp += (double)w[k - 1] * v[k];
   if (act) __builtin_nontemporal_store(tmp * a.scale, a.dVm + i);
}

__global__ __launch_bounds__(256) void k_integrate_f64(int n, double dt, double *__restrict__ Vm,
                                                       const double *__restrict__ dVmR,
                                                       const double *__restrict__ dVmD,
                                                       const double *__restrict__ iStim)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) Vm[i] += dt * (dVmR[i] + dVmD[i] + iStim[i]);
}

__global__ __launch_bounds__(256) void k_fill_f64(int n, double v, double *__restrict__ x)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) x[i] = v;
}

__global__ __launch_bounds__(256) void k_set_list_f64(int n, const int *__restrict__ list, double v,
                                                      double *__restrict__ x)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) x[list[i]] = v;
}

// Sensor evaluation at (time, loop) after the step; Vm in cell order.
__global__ __launch_bounds__(256) void k_sensor_f64(SensorArgs a, const double *__restrict__ Vm,
                                                    const double *__restrict__ vrest)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   const uint32_t ev = i < a.n ? sensor_events_packed(sensor_eval(a, i, Vm[i], a.apd ? vrest[i] : 0.0)) : 0u;
   sensor_count(a.counts, ev);
}

// The same, four cells per thread: the flags of the four as one 4-byte
// load each, Vm as two 16-byte loads (Vm 16-byte aligned; the host checks).
__global__ __launch_bounds__(256) void k_sensor4_f64(SensorArgs a, const double *__restrict__ Vm,
                                                     const double *__restrict__ vrest)
{
   const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
   uint32_t ev = 0;
   if (i + 3 < a.n) {
      const uint32_t acd = *reinterpret_cast<const uint32_t *>(a.activated + i);
      const uint32_t act = *reinterpret_cast<const uint32_t *>(a.active + i);
      const double2 v01 = *reinterpret_cast<const double2 *>(Vm + i);
      const double2 v23 = *reinterpret_cast<const double2 *>(Vm + i + 2);
      const double v[4] = {v01.x, v01.y, v23.x, v23.y};
#pragma unroll
      for (int k = 0; k < 4; k++)
         ev += sensor_events_packed(sensor_eval_flags(a, (int)i + k, v[k],
                                                      a.apd ? vrest[i + k] : 0.0,
                                                      (uint8_t)(acd >> (8 * k)),
                                                      (uint8_t)(act >> (8 * k))));
   } else {
      for (int k = (int)i; k < a.n; k++)
         ev += sensor_events_packed(sensor_eval(a, k, Vm[k], a.apd ? vrest[k] : 0.0));
   }
   sensor_count(a.counts, ev);
}

// Simulate::checkRanges on Vm before the integrator (elec/simulationLoop.cc:214-217).
__global__ __launch_bounds__(256) void k_check_range_f64(int n, const double *__restrict__ Vm,
                                                         const double *__restrict__ dVmR,
                                                         const double *__restrict__ dVmD,
                                                         double vmin, double vmax, int loop,
                                                         RangeRec *r)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   const double v = i < n ? Vm[i] : 0.0;
   const bool out = i < n && (v > vmax || v < vmin);
   const uint64_t m = __ballot(out);
   if (m) range_report(r, m, threadIdx.x & 63, loop, i, v, out ? dVmR[i] : 0.0, out ? dVmD[i] : 0.0);
}

}  // namespace cdh
