// tt06_mixed.hpp -- mixed-precision TT06Dev cell update for gfx950 (device code).
//
// Storage: Vm, Ca_i and Na_i in FP64; the other 17 states in FP32, with K_i
// held as its offset from the cell type's rest K_i (an FP32 ulp of K_i itself
// would be ~1.5e-5 mM, the size of one step's change).  Arithmetic: the
// rational-fit polynomials run as FP64 fused multiply-adds (FP32 Horner
// loses up to 3e-2 of a fit's range, and a non-packed FP32 FMA issues at the
// FP64 rate on this part); their ratios and the other divisions use the
// hardware FP64 reciprocal, the four logs the FP32 hardware log.  Everything else is FP64 with
// contraction, rounded to FP32 on store.
//
// Same model as tt06_f64.hpp (elec/TT06NonGates.c:374-563, the v2 gates of
// elec/TT06GatesSimd.c), but not bitwise: the gate updates all use the
// fused form g + r*(mhu - g).
#pragma once

#include <hip/hip_runtime.h>

#include "tt06_f64.hpp"
#include "tt06_model.hpp"

namespace cdh {
namespace mixed {

// FP32 state slots
enum Slot32 : int {
   M_CA_SS = 0, M_CA_SR, M_R_PRIME, M_FCASS, M_DK_I,
   M_M, M_H, M_J, M_XR1, M_XR2, M_XS, M_R, M_D, M_F, M_F2, M_JL, M_S,
   N_STATE32
};
enum Slot64 : int { M_CA_I = 0, M_NA_I, N_STATE64 };

// Horner with fused multiply-adds (tt06_f64.hpp)
template <int LO, int HI>
__device__ __forceinline__ double hornerFma(const double *__restrict__ a, double x)
{
   return horner64<LO, HI, true>(a, x);
}

__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }
// num/den through the hardware FP64 reciprocal (v_rcp_f64, measured on this
// part within 4.6e-8 relative; the FP32 reciprocal of the rounded
// denominator is within 1.2e-7) and one multiply: two instructions where
// the FP32 route needs three conversions besides
__device__ __forceinline__ double ratio(double num, double den)
{
   return num * __builtin_amdgcn_rcp(den);
}
__device__ __forceinline__ double log32(double x) { return (double)__logf((float)x); }

template <int I>
__device__ __forceinline__ double fitM(const double *__restrict__ coef, double x)
{
   constexpr int off = fit_offset(I);
   constexpr int l = kFitOrder[I].l, m = kFitOrder[I].m;
   const double num = hornerFma<0, m - 1>(coef + off, x);
   if constexpr (l < 2) {
      return num;
   } else {
      return ratio(num, hornerFma<m, m + l - 1>(coef + off, x));
   }
}

// one gate: g + r*(mhu - g); r = 1 - 1/T16(x) for m (Rush-Larsen with the
// reference's degree-16 series), r = dt/tau for the rest
template <int EQ, bool RL>
__device__ __forceinline__ float gateM(const double *__restrict__ coef, const double *__restrict__ expc,
                                       double V, float g)
{
   const float mhu = (float)fitM<7 + 2 * EQ>(coef, V);
   const double x = fitM<8 + 2 * EQ>(coef, V);
   float r;
   if constexpr (RL) {
      const double e = hornerFma<0, 16>(expc, x);
      r = 1.0f - rcp32((float)e);
   } else {
      r = (float)x;
   }
   return __builtin_fmaf(r, mhu - g, g);
}

__device__ __forceinline__ double sigmM(double x) { return ratio(x, 1.0 + x); }

// Non-gate update.  s32/s64 are the cell's states; dK0 is the rest K_i of its
// cell type.  Returns dVdt = -Iion and the potassium current iK (the
// caller finishes the K_i update, which needs the step's diffusion and
// stimulus currents).
__device__ __forceinline__ double nongateM(const TT06Const &K, const ParamSet &P,
                                           const double *__restrict__ coef, double dt, double Vm,
                                           const double K_i, float *s32, double *s64, double *iK_out)
{
   const double fv0 = fitM<0>(coef, Vm), fv1 = fitM<1>(coef, Vm), fv2 = fitM<2>(coef, Vm);
   const double fv3 = fitM<3>(coef, Vm), fv4 = fitM<4>(coef, Vm), fv5 = fitM<5>(coef, Vm);
   const double Na_i = s64[M_NA_I], Ca_i = s64[M_CA_I];
   const double Ca_ss = s32[M_CA_SS], Ca_SR = s32[M_CA_SR];
   const double fCass = s32[M_FCASS], R_prime = s32[M_R_PRIME];
   const double dtc9 = dt * K.c9;

   // calcium
   const double x1 = Ca_i * K.c26;
   const double a2 = Ca_i * K.c27, x2 = a2 * a2;
   const double a3 = Ca_i * K.c28 + K.c29, x3 = a3 * a3;
   const double sg1 = sigmM(x1), sg2 = sigmM(x2), sg3 = sigmM(x3);
   const double dV3 = Vm - 0.5 * K.c3 * log32(Ca_i) - K.c8;
   const double I_NaCa = (Na_i * Na_i * Na_i) * fv1 - Ca_i * fv2;
   const double I_bCa = K.c7 * dV3;
   const double I_pCa = K.c24 * sg1;
   const double I_up = K.c44 * sg2;
   const double I_leak = K.c43 * (Ca_SR - Ca_i);
   const double I_xfer = K.c23 * (Ca_ss - Ca_i);
   const double I_delta = I_leak - I_up;
   const double I_sum = I_bCa + I_pCa;
   s64[M_CA_I] = Ca_i + dtc9 * (sg3 * (0.5 * I_sum - I_NaCa + I_xfer * K.c15 + I_delta * K.c16));
   double dVR = -I_sum;

   // potassium and sodium
   const double lK = log32(K_i);
   const double dV0 = Vm - K.c3 * lK - K.c5;
   const double dV1 = Vm - K.c3 * log32(Na_i) - K.c4;
   const double dV2 = dV0 - K.c3 * log32(1 + ratio(K.c2 * Na_i, K_i)) + K.c5 - K.c6;
   const double fv6 = fitM<6>(coef, dV0);
   const double m = s32[M_M], h = s32[M_H], j = s32[M_J];
   const double m3 = m * m * m;
   const double I_NaK = P.P_NaK * sigmM(Na_i * K.c25) * fv0;
   const double I_pK = fv5 * dV0;
   const double I_K1 = fv6 * dV0;
   const double I_to = P.g_to * (double)s32[M_R] * (double)s32[M_S] * dV0;
   const double I_Kr = P.c_Kr * (double)s32[M_XR1] * (double)s32[M_XR2] * dV0;
   const double I_Na = P.c_Na * m3 * h * j * dV1;
   const double I_NaL = P.g_NaL * m3 * (double)s32[M_JL] * dV1;
   const double I_bNa = K.c21 * dV1;
   const double xs = s32[M_XS];
   const double I_Ks = P.g_Ks * (xs * xs) * dV2;
   const double iNa = 3 * I_NaCa - 1.5 * I_NaK + I_Na + I_bNa + I_NaL;
   const double iK = I_Ks + I_NaK + I_pK + I_K1 + I_to + I_Kr;
   dVR -= iNa + iK - 2.0 * I_NaCa;
   s64[M_NA_I] = Na_i + dtc9 * iNa;

   // Ca_ss, Ca_SR, R', fCass
   const double a4 = Ca_SR * K.c30, x4 = a4 * a4;
   const double a5 = Ca_SR * K.c31 + K.c32, x5 = a5 * a5;
   const double a6 = Ca_ss * K.c33 + K.c34, x6 = a6 * a6;
   const double sg4 = sigmM(x4), sg5 = sigmM(x5), sg6 = sigmM(x6);
   const double tmp8 = K.c18 + K.c19 * sg4;
   const double tmp9 = tmp8 * Ca_ss + K.c36;
   const double I_CaL = P.c_CaL * (double)s32[M_D] * (double)s32[M_F] * (double)s32[M_F2] * fCass *
                        (fv4 - Ca_ss * fv3);
   const double css2 = Ca_ss * Ca_ss;
   const double O = ratio(css2 * R_prime, tmp8 * K.c17 + css2);
   const double I_rel = K.c40 * O * (Ca_SR - Ca_ss);
   s32[M_CA_SS] = (float)(Ca_ss + dtc9 * sg6 * (I_xfer + I_rel * K.c14 + I_CaL * K.c13));
   s32[M_CA_SR] = (float)(Ca_SR - dtc9 * sg5 * (I_delta + I_rel));
   s32[M_R_PRIME] = (float)(R_prime + dtc9 * (K.c36 - tmp9 * R_prime));
   const double z = 20 * Ca_ss;
   const float t1 = rcp32((float)(1.0 + z * z));
   const float mhu = 0.6f * t1 + 0.4f;
   const float tauR = rcp32(80.0f * t1 + 2.0f);
   s32[M_FCASS] = (float)(fCass + dt * (mhu - fCass) * tauR);
   dVR += I_CaL;
   *iK_out = iK;
   return dVR;
}

__device__ __forceinline__ void gatesM(const double *__restrict__ coef, const double *__restrict__ expc,
                                       int s_switch, uint8_t flags, double Vm, float *s32,
                                       uint32_t gmask)
{
   if (flags & kFlagGatesFixed) return;
   if (gmask & (1u << 0)) s32[M_M] = gateM<0, true>(coef, expc, Vm, s32[M_M]);
   if (gmask & (1u << 1)) s32[M_H] = gateM<1, false>(coef, expc, Vm, s32[M_H]);
   if (gmask & (1u << 2)) s32[M_J] = gateM<2, false>(coef, expc, Vm, s32[M_J]);
   if (gmask & (1u << 3)) s32[M_XR1] = gateM<3, false>(coef, expc, Vm, s32[M_XR1]);
   if (gmask & (1u << 4)) s32[M_XR2] = gateM<4, false>(coef, expc, Vm, s32[M_XR2]);
   if (gmask & (1u << 5)) s32[M_XS] = gateM<5, false>(coef, expc, Vm, s32[M_XS]);
   if (gmask & (1u << 6)) s32[M_R] = gateM<6, false>(coef, expc, Vm, s32[M_R]);
   if (gmask & (1u << 7)) s32[M_D] = gateM<7, false>(coef, expc, Vm, s32[M_D]);
   if (gmask & (1u << 8)) s32[M_F] = gateM<8, false>(coef, expc, Vm, s32[M_F]);
   if (gmask & (1u << 9)) s32[M_F2] = gateM<9, false>(coef, expc, Vm, s32[M_F2]);
   if (gmask & (1u << 10)) s32[M_JL] = gateM<10, false>(coef, expc, Vm, s32[M_JL]);
   if (gmask & (1u << 11)) {
      if (s_switch) s32[M_S] = gateM<12, false>(coef, expc, Vm, s32[M_S]);
      else          s32[M_S] = gateM<11, false>(coef, expc, Vm, s32[M_S]);
   }
}

}  // namespace mixed
}  // namespace cdh
