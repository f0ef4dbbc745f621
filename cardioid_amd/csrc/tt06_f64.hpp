// tt06_f64.hpp -- FP64 TT06Dev cell update for gfx950 (device code).
//
// One cell per lane.  Every expression keeps the reference's operation
// order, and the translation units that include this file are compiled
// with -ffp-contract=off, so each multiply and add rounds on its own as in
// the reference's x86 build.  The results are then bitwise those of
//   update_nonGate            elec/TT06NonGates.c:374-563
//   update_*Gate_v2           elec/TT06GatesSimd.c (update forms per gate)
// except where the device log() and glibc's log() round differently.
#pragma once

#include <hip/hip_runtime.h>

#include "tt06_model.hpp"

namespace cdh {

// Per parameter-set values the kernels read (one entry per cell type, or
// per (cell type, drug scales) pair).  c_Kr = s_Kr*g_Kr, c_Na = s_Na*c20 and
// c_CaL = s_CaL are formed on the host with the reference's products
// (TT06NonGates.c:405-426), so unit scales give the reference's doubles.
struct ParamSet {
   double P_NaK, g_Ks, c_Kr, g_to, g_NaL, c_CaL, c_Na;
   double K0;        // rest K_i of the cell type (the mixed form stores K_i - K0)
   double Vm0;       // rest Vm of the cell type (APD90's Vrest)
   int s_switch;
   int pad_;
};

constexpr int kMaxParamSets = 256;

// x*s + a with the coefficient a taken from SGPRs by the VOP3 encoding.
// Written out because the compiler otherwise picks v_fmac_f64 (addend tied
// to the destination) and copies every coefficient into VGPRs first, two
// v_mov per Horner step.
__device__ __forceinline__ double fma_vvs(double x, double s, double a)
{
   double r;
   asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(s), "s"(a));
   return r;
}

// x*a and x + a with a taken from SGPRs (VOP3, one scalar operand), written
// out so that no contraction setting fuses them into an FMA whose two
// coefficients would both need SGPRs
__device__ __forceinline__ double mul_vs(double x, double a)
{
   double r;
   asm("v_mul_f64 %0, %1, %2" : "=v"(r) : "v"(x), "s"(a));
   return r;
}
__device__ __forceinline__ double add_vs(double x, double a)
{
   double r;
   asm("v_add_f64 %0, %1, %2" : "=v"(r) : "v"(x), "s"(a));
   return r;
}

// Horner of a[LO..HI] at x.  FMA = false: separate multiply and add, the
// reference's rounding (pade.cc:17-36 built without contraction).  FMA =
// true: fused multiply-adds; the first step stays a multiply and an add,
// which take their coefficient straight from SGPRs (a first FMA would need
// one of its two coefficients copied into VGPRs: one scalar operand per
// instruction).
template <int LO, int HI, bool FMA = false, class P>
__device__ __forceinline__ double horner64(P a, double x)
{
   if constexpr (FMA && HI > LO) {
      double s = add_vs(mul_vs(x, a[HI]), a[HI - 1]);
#pragma unroll
      for (int j = HI - 2; j >= LO; --j) s = fma_vvs(x, s, a[j]);
      return s;
   } else {
      double s = a[HI];
#pragma unroll
      for (int j = HI - 1; j >= LO; --j) s = a[j] + x * s;
      return s;
   }
}

// rational fit i of the fixed layout (pade.cc:17-36: num/den, polynomial if l < 2)
template <int I, bool FMA = false>
__device__ __forceinline__ double fit64(const double *__restrict__ coef, double x)
{
   constexpr int off = fit_offset(I);
   constexpr int l = kFitOrder[I].l, m = kFitOrder[I].m;
   const double *__restrict__ c = coef + off;
   const double num = horner64<0, m - 1, FMA>(c, x);
   if constexpr (l < 2) {
      return num;
   } else {
      const double den = horner64<m, m + l - 1, FMA>(c, x);
      return num / den;
   }
}

enum GateForm { G_RL, G_DIF, G_SUM };

// one gate: mhu and dt/tau from fits 7+2*eq and 8+2*eq, then the v2 update
template <int EQ, GateForm FORM, bool FMA = false>
__device__ __forceinline__ double gate64(const double *__restrict__ coef, const double *__restrict__ expc,
                                         double V, double g, bool sub_form)
{
   const double mhu = fit64<7 + 2 * EQ, FMA>(coef, V);
   const double x = fit64<8 + 2 * EQ, FMA>(coef, V);
   if constexpr (FORM == G_RL) {
      const double e = horner64<0, 16, FMA>(expc, x);
      const double r = 1.0 - 1.0 / e;
      return r * (mhu - g) + g;
   } else if constexpr (FORM == G_SUM) {
      return (mhu * x + g) - g * x;
   } else {
      if (sub_form) return (mhu - g) * x + g;   // f gate, main-loop lanes 0..3
      return x * mhu + (g - x * g);
   }
}

#define CDH_SIGM(x) ((x) / (1.0 + (x)))

// The seven voltage functions of the non-gate update (fv05General /
// fv6General, elec/TT06Func.cc:614-629) from the dt-independent rational
// fits 0..6 (configuration B).  The exact functions of configuration A are
// the other implementation of the same two calls (tt06_exact.hpp).
template <bool FMA>
struct FitVoltageFns {
   const double *__restrict__ coef;
   __device__ __forceinline__ void fv05(double Vm, double fv[6]) const
   {
      fv[0] = fit64<0, FMA>(coef, Vm);
      fv[1] = fit64<1, FMA>(coef, Vm);
      fv[2] = fit64<2, FMA>(coef, Vm);
      fv[3] = fit64<3, FMA>(coef, Vm);
      fv[4] = fit64<4, FMA>(coef, Vm);
      fv[5] = fit64<5, FMA>(coef, Vm);
   }
   __device__ __forceinline__ double fv6(double dV0) const { return fit64<6, FMA>(coef, dV0); }
};

// Non-gate update of one cell (update_nonGate, elec/TT06NonGates.c:374-563);
// returns dVdt = -Iion and updates the seven concentration-like states in
// st[].  Gates are read, not written.  FV supplies fv0..fv6.
template <class FV>
__device__ __forceinline__ double nongate_core64(const TT06Const &K, const ParamSet &P,
                                                 const FV &fvs, double dt, double Vm, double *st)
{
   double fvv[6];
   fvs.fv05(Vm, fvv);
   const double fv0 = fvv[0], fv1 = fvv[1], fv2 = fvv[2], fv3 = fvv[3], fv4 = fvv[4];
   const double fv5 = fvv[5];
   const double Na_i = st[S_NA_I], Ca_i = st[S_CA_I], Ca_ss = st[S_CA_SS], Ca_SR = st[S_CA_SR];
   const double fCass = st[S_FCASS], R_prime = st[S_R_PRIME];
   double dVK_i = st[S_DVK_I];
   const double c_K1 = 1, c_bNa = 1 * K.c21, c_bCa = 1 * K.c7, c_NaCa = 1, c_pCa = 1 * K.c24;
   const double c_pK = 1;
   const double c_NaK = 1 * P.P_NaK, c_Ks = 1 * P.g_Ks, c_to = 1 * P.g_to, c_NaL = 1 * P.g_NaL;

   double dVR = 0.0;
   // calcium
   const double x1 = Ca_i * K.c26;
   const double x2 = (Ca_i * K.c27) * (Ca_i * K.c27);
   const double x3 = (Ca_i * K.c28 + K.c29) * (Ca_i * K.c28 + K.c29);
   const double sg1 = CDH_SIGM(x1), sg2 = CDH_SIGM(x2), sg3 = CDH_SIGM(x3);
   const double dV3 = Vm - 0.5 * K.c3 * log(Ca_i) - K.c8;
   const double I_NaCa = c_NaCa * ((Na_i * Na_i * Na_i) * fv1 - Ca_i * fv2);
   const double I_bCa = c_bCa * dV3;
   const double I_pCa = c_pCa * sg1;
   const double I_up = K.c44 * sg2;
   const double I_leak = K.c43 * (Ca_SR - Ca_i);
   const double I_xfer = K.c23 * (Ca_ss - Ca_i);
   const double I_delta = I_leak - I_up;
   const double I_sum = I_bCa + I_pCa;
   st[S_CA_I] = Ca_i + (dt * K.c9) * (sg3 * (0.5 * I_sum - I_NaCa + I_xfer * K.c15 + I_delta * K.c16));
   dVR -= I_sum;

   // potassium and sodium
   const double K_i = K.c9 * (dVK_i - Vm);
   const double x0 = Na_i * K.c25;
   const double dV0 = Vm - K.c3 * log(K_i) - K.c5;
   const double dV1 = Vm - K.c3 * log(Na_i) - K.c4;
   const double dV2 = dV0 - K.c3 * log(1 + K.c2 * Na_i / K_i) + K.c5 - K.c6;
   const double fv6 = fvs.fv6(dV0);
   const double m = st[S_M], h = st[S_H], j = st[S_J];
   const double I_NaK = c_NaK * CDH_SIGM(x0) * fv0;
   const double I_pK = c_pK * fv5 * dV0;
   const double I_K1 = c_K1 * fv6 * dV0;
   const double I_to = c_to * st[S_R] * st[S_S] * dV0;
   const double I_Kr = P.c_Kr * st[S_XR1] * st[S_XR2] * dV0;
   const double I_Na = P.c_Na * (m * m * m) * h * j * dV1;
   const double I_NaL = c_NaL * (m * m * m) * st[S_JL] * dV1;
   const double I_bNa = c_bNa * dV1;
   const double I_Ks = c_Ks * (st[S_XS] * st[S_XS]) * dV2;
   const double iNa = 3 * I_NaCa - 1.5 * I_NaK + I_Na + I_bNa + I_NaL;
   const double iK = I_Ks + I_NaK + I_pK + I_K1 + I_to + I_Kr;
   dVK_i += dt * iK;
   dVR -= iNa + iK - 2.0 * I_NaCa;
   st[S_NA_I] = Na_i + (dt * K.c9) * iNa;

   // Ca_ss, Ca_SR, R', fCass
   const double x4 = (Ca_SR * K.c30) * (Ca_SR * K.c30);
   const double x5 = (Ca_SR * K.c31 + K.c32) * (Ca_SR * K.c31 + K.c32);
   const double x6 = (Ca_ss * K.c33 + K.c34) * (Ca_ss * K.c33 + K.c34);
   const double sg4 = CDH_SIGM(x4), sg5 = CDH_SIGM(x5), sg6 = CDH_SIGM(x6);
   const double tmp8 = (K.c18 + K.c19 * sg4);
   const double tmp9 = tmp8 * Ca_ss + K.c36;
   const double I_CaL = P.c_CaL * st[S_D] * st[S_F] * st[S_F2] * fCass * (fv4 - Ca_ss * fv3);
   const double O = (Ca_ss * Ca_ss) * R_prime / (tmp8 * K.c17 + (Ca_ss * Ca_ss));
   const double I_rel = K.c40 * O * (Ca_SR - Ca_ss);
   st[S_CA_SS] = Ca_ss + (dt * K.c9) * sg6 * (I_xfer + I_rel * K.c14 + I_CaL * K.c13);
   st[S_CA_SR] = Ca_SR - (dt * K.c9) * sg5 * (I_delta + I_rel);
   st[S_R_PRIME] = R_prime + (dt * K.c9) * (K.c36 - tmp9 * R_prime);
   const double t1 = 1.0 / (1.0 + (20 * Ca_ss) * (20 * Ca_ss));
   const double mhu = 0.600000 * t1 + 0.4000000;
   const double tauR = 1.0 / (80.0 * t1 + 2.0);
   st[S_FCASS] = fCass + dt * (mhu - fCass) * tauR;
   dVR += I_CaL;
   st[S_DVK_I] = dVK_i + dt * dVR;
   return dVR;
}

// configuration B: the fits (FMA applies to the fit polynomials only)
template <bool FMA = false>
__device__ __forceinline__ double nongate64(const TT06Const &K, const ParamSet &P,
                                           const double *__restrict__ coef, double dt, double Vm,
                                           double *st)
{
   return nongate_core64(K, P, FitVoltageFns<FMA>{coef}, dt, Vm, st);
}

// The 12 gate updates at voltage Vm (the voltage of the step's start).
//
// `gmask` is a uniform kernel argument with every bit set.  Testing it puts
// each gate in a basic block of its own, which keeps the compiler from
// issuing the scalar loads of all 26 gate fits ahead of time (they would
// not fit in SGPRs and get spilled to VGPR lanes).
template <bool FMA = false>
__device__ __forceinline__ void gates64(const double *__restrict__ coef, const double *__restrict__ expc,
                                        int s_switch, uint8_t flags, double Vm, double *st,
                                        uint32_t gmask)
{
   if (flags & kFlagGatesFixed) return;
   const bool sub = (flags & kFlagFGateSub) != 0;
   if (gmask & (1u << 0)) st[S_M] = gate64<0, G_RL, FMA>(coef, expc, Vm, st[S_M], false);
   if (gmask & (1u << 1)) st[S_H] = gate64<1, G_DIF, FMA>(coef, expc, Vm, st[S_H], false);
   if (gmask & (1u << 2)) st[S_J] = gate64<2, G_DIF, FMA>(coef, expc, Vm, st[S_J], false);
   if (gmask & (1u << 3)) st[S_XR1] = gate64<3, G_DIF, FMA>(coef, expc, Vm, st[S_XR1], false);
   if (gmask & (1u << 4)) st[S_XR2] = gate64<4, G_SUM, FMA>(coef, expc, Vm, st[S_XR2], false);
   if (gmask & (1u << 5)) st[S_XS] = gate64<5, G_DIF, FMA>(coef, expc, Vm, st[S_XS], false);
   if (gmask & (1u << 6)) st[S_R] = gate64<6, G_DIF, FMA>(coef, expc, Vm, st[S_R], false);
   if (gmask & (1u << 7)) st[S_D] = gate64<7, G_DIF, FMA>(coef, expc, Vm, st[S_D], false);
   if (gmask & (1u << 8)) st[S_F] = gate64<8, G_DIF, FMA>(coef, expc, Vm, st[S_F], sub);
   if (gmask & (1u << 9)) st[S_F2] = gate64<9, G_DIF, FMA>(coef, expc, Vm, st[S_F2], false);
   if (gmask & (1u << 10)) st[S_JL] = gate64<10, G_DIF, FMA>(coef, expc, Vm, st[S_JL], false);
   if (gmask & (1u << 11)) {
      if (s_switch) st[S_S] = gate64<12, G_DIF, FMA>(coef, expc, Vm, st[S_S], false);
      else          st[S_S] = gate64<11, G_DIF, FMA>(coef, expc, Vm, st[S_S], false);
   }
}

}  // namespace cdh
