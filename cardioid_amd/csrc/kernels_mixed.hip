// kernels_mixed.hip -- the mixed-precision fused step (gfx950).
//
// Same step as k_step_fused_f64 (elec/simulationLoop.cc:170-241 for one
// cell), with the storage and arithmetic of tt06_mixed.hpp and an FP32
// stencil on voltage differences:
//    dVm_D = scale * sum_{k=1..18} A_k * (phi_k - phi_0)
// which is the FGR operator itself (its centre weight is minus the sum of
// the others, elec/FGRDiffusionOMP.cc:214-219) without the centre weight.
#pragma clang fp contract(fast)

#include "kernels.hpp"
#include "tt06_mixed.hpp"

namespace cdh {

using namespace mixed;

template <int WAVES>
__global__ __launch_bounds__(256, WAVES) void k_step_fused_mixed(MixedArgs a)
{
   const int i = xcd_cell_index(a.n, a.nblock_per_xcd);
   if (i >= a.n) return;
   const uint32_t b = a.bi[i];
   const double *phi = a.box_cur + b;
   const double V = phi[0];

   // states and weights stream once per step: 64-cell blocks, non-temporal
   float *p32 = a.st32 + blk_base(i, N_STATE32);
   double *p64 = a.st64 + blk_base(i, N_STATE64);
   float s32[N_STATE32];
#pragma unroll
   for (int s = 0; s < N_STATE32; s++) s32[s] = __builtin_nontemporal_load(p32 + kBlk * s);
   double s64[N_STATE64];
#pragma unroll
   for (int s = 0; s < N_STATE64; s++) s64[s] = __builtin_nontemporal_load(p64 + kBlk * s);
   const ParamSet P = a.params[a.pid[i]];
   const double K_i = P.K0 + (double)s32[M_DK_I];
   double iK;
   const double dVR = nongateM(a.K, P, a.coef, a.dt, V, K_i, s32, s64, &iK);
   gatesM(a.coef, a.K.exp_coef, P.s_switch, a.flags[i], V, s32, a.gate_mask);

   const float *wp = a.Aw + blk_base(i, 18);
   float acc = 0.0f;
#pragma unroll
   for (int k = 1; k < 19; k++)
      acc = __builtin_fmaf(__builtin_nontemporal_load(wp + kBlk * (k - 1)),
                           (float)(phi[a.off[k]] - V), acc);
   const double dVD = (double)acc * a.scale;

   double iStim = 0.0;
   if (a.stim_any) iStim = stim_value(a.stim, a.stim_mask[i]);

   const double Vn = V + a.dt * (dVR + dVD + iStim);
   // K_i follows dVK_i = K_i/c9 + Vm of the reference: dK_i = c9*dt*(iK - dVm_D - iStim)
   s32[M_DK_I] = (float)((double)s32[M_DK_I] + (a.dt * a.K.c9) * (iK - dVD - iStim));
#pragma unroll
   for (int s = 0; s < N_STATE32; s++) __builtin_nontemporal_store(s32[s], p32 + kBlk * s);
#pragma unroll
   for (int s = 0; s < N_STATE64; s++) __builtin_nontemporal_store(s64[s], p64 + kBlk * s);
   a.box_next[b] = Vn;
   if (a.sens.enabled) sensor_eval_pair(a.sens, i, V, Vn, P.Vm0);
}

template __global__ void k_step_fused_mixed<3>(MixedArgs);
template __global__ void k_step_fused_mixed<4>(MixedArgs);
template __global__ void k_step_fused_mixed<5>(MixedArgs);
template __global__ void k_step_fused_mixed<6>(MixedArgs);

}  // namespace cdh
