// kernels_f64.hip -- FP64 kernels of the drop-in path (gfx950).
//
// Compiled with -ffp-contract=off: every kernel here reproduces the
// reference CPU path's rounding, operation by operation.
//
//   k_tt06_calc_f64     TT06Dev_Reaction::calc          elec/TT06Dev_Reaction.cc:393-407
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
#include "tt06_exact.hpp"
#include "tt06_f64.hpp"

namespace cdh {

// The drop-in objects keep their states and weights in 64-cell tiles:
// field q of cell i at blk_index(q, i, nq), a wavefront's fields one
// contiguous run.
__global__ __launch_bounds__(256) void k_tt06_calc_f64(ReactionArgs a)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= a.n) return;
   const int e = a.index ? a.index[i] : i;
   const double V = a.Vm[e];
   double st[N_STATE];
#pragma unroll
   for (int s = 0; s < N_STATE; s++) st[s] = a.state[blk_index(s, i, N_STATE)];
   const ParamSet P = a.params[a.pid ? a.pid[i] : 0];
   // the non-gate update in a basic block of its own (gate_mask bit 12,
   // always set): otherwise the compiler hoists its scalar coefficient loads
   // to the kernel entry and spills them
   double dVR = 0.0;
   if (a.gate_mask & (1u << 12)) dVR = nongate64(*a.K, P, a.coef, a.dt, V, st);
   gates64(a.coef, a.K->exp_coef, P.s_switch, a.flags[i], V, st, a.gate_mask);
#pragma unroll
   for (int s = 0; s < N_STATE; s++) a.state[blk_index(s, i, N_STATE)] = st[s];
   a.dVm[e] = dVR;
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

// Blocks are dealt round-robin to the 8 XCDs; each XCD takes a contiguous
// range of cells, so that its L2 holds the box rows and planes they read.
__global__ __launch_bounds__(256) void k_fgr_calc_f64(StencilArgs a)
{
   const int xcd = blockIdx.x & 7;
   const int64_t i =
       ((int64_t)xcd * a.groups_per_xcd + (blockIdx.x >> 3)) * blockDim.x + threadIdx.x;
   if (i >= a.n) return;
   // Every load is issued before any arithmetic: the box index, then the
   // weights (they stream while the index arrives), then the 19 neighbours.
   // Left to itself the compiler interleaves the neighbour loads with waits
   // on the weights for the centre-weight sum, which costs a third round
   // trip to memory per wave (4 % of the kernel's time).
   const uint32_t b = a.bi[i];
   float w[18];
#pragma unroll
   for (int k = 0; k < 18; k++) w[k] = __builtin_nontemporal_load(a.Aw + blk_index(k, i, 18));
   __builtin_amdgcn_sched_barrier(0);
   const double *phi = a.box + b;
   double v[19];
   v[0] = phi[0];
#pragma unroll
   for (int k = 1; k < 19; k++) v[k] = phi[a.off[k]];
   __builtin_amdgcn_sched_barrier(0);
   // the weights stream once per step; the centre weight is minus their sum
   // in order, in double (FGRDiffusionOMP.cc:214-219), recomputed bit for bit
   double A0 = 0.0;
#pragma unroll
   for (int k = 0; k < 18; k++) A0 -= w[k];
   double tmp = A0 * v[0];
#pragma unroll
   for (int k = 1; k < 19; k++) tmp += (double)w[k - 1] * v[k];
   __builtin_nontemporal_store(tmp * a.scale, a.dVm + i);
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

// Integrator fused with the activation sensors (pair form: Vm before and
// after the step are both at hand here).
__global__ __launch_bounds__(256) void k_integrate_sensor_f64(int n, double dt,
                                                              double *__restrict__ Vm,
                                                              const double *__restrict__ dVmR,
                                                              const double *__restrict__ dVmD,
                                                              const double *__restrict__ iStim,
                                                              SensorArgs s,
                                                              const ParamSet *__restrict__ params,
                                                              const uint8_t *__restrict__ pid)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const double v0 = Vm[i];
   const double v1 = v0 + dt * (dVmR[i] + dVmD[i] + iStim[i]);
   Vm[i] = v1;
   if (s.enabled) sensor_eval_pair(s, i, v0, v1, s.apd ? params[pid[i]].Vm0 : 0.0);
}

// Sensor evaluation at (time, loop) after the step; Vm in cell order.
__global__ __launch_bounds__(256) void k_sensor_f64(SensorArgs a, const double *__restrict__ Vm,
                                                    const double *__restrict__ vrest)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= a.n) return;
   sensor_eval(a, i, Vm[i], a.apd ? vrest[i] : 0.0);
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
