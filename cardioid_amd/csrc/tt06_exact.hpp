// tt06_exact.hpp -- configuration A of the TT06Dev reaction on gfx950: the
// exact TT06 voltage functions instead of rational fits.
//
// REACTION_FACTORY(TT06Dev) defaults to tolerance = 0 (elec/TT06Dev_Reaction.cc:600-602);
// then padeApprox keeps each function itself (elec/pade.cc:237-268), the
// reaction runs the generic gate update TT06Func::updateGate
// (elec/TT06Func.cc:64-161: Rush-Larsen with exp for m, dt*(mhu - g)*tauR for
// the others, every cell of the group) and the generic non-gate update with
// fv0..fv6 (elec/TT06NonGates.c:319-364).  mod = 1 (jhTauSmooth) swaps the h
// and j rates for their smoothed forms (TauRecipMod, elec/TT06Tau.cc:131-151):
// a cubic Hermite bridge on [x0, x1], whose coefficients the host computes.
//
// Every expression keeps the reference's operation order and the including
// translation unit is compiled without contraction, so results differ from
// the reference's x86 build only where the device exp / log / sqrt / pow
// round differently from glibc's (a few ulps).
#pragma once

#include <hip/hip_runtime.h>

#include "tt06_f64.hpp"

namespace cdh {

namespace exact {

__device__ __forceinline__ double sq(double x) { return x * x; }

// gate steady states and rates, elec/TT06Func.cc:261-417
__device__ __forceinline__ double mMhu(double V) { return 1.0 / sq(1.0 + exp((-56.8600 - V) / 9.03000)); }
__device__ __forceinline__ double mTauR(double V)
{
   const double t1 = 1.0 / (1.0 + (exp(((-60.0 - V) / 5.0))));
   const double t2 = 0.10000 / (1.0 + (exp(((V + 35.0) / 5.0)))) +
                     0.100000 / (1.0 + (exp(((V - 50.0) / 200.0))));
   return 1.0 / (t1 * t2);
}
__device__ __forceinline__ double hjMhu(double V) { return 1.0 / sq((1.0 + (exp(((V + 71.5500) / 7.43000))))); }
__device__ __forceinline__ double hTauR(double V)
{
   const double t1 = (V < -40.0 ? 0.0570000 * (exp((-(V + 80.0) / 6.80000))) : 0.0);
   const double t2 = (V < -40.0 ? 2.70000 * (exp((0.0790000 * V))) + 310000. * (exp((0.348500 * V)))
                                : 0.770000 / (0.130000 * (1.0 + (exp(((V + 10.6600) / -11.1000))))));
   return t1 + t2;
}
__device__ __forceinline__ double jTauR(double V)
{
   const double t1 = (V < -40.0 ? (((-25428.0 * (exp((0.244400 * V))) - 6.94800e-06 * (exp((-0.0439100 * V)))) *
                                   (V + 37.7800)) / 1.0) / (1.0 + (exp((0.311000 * (V + 79.2300)))))
                                : 0.0);
   const double t2 = (V < -40.0 ? (0.0242400 * (exp((-0.0105200 * V)))) / (1.0 + (exp((-0.137800 * (V + 40.1400)))))
                                : (0.600000 * (exp((0.0570000 * V)))) / (1.0 + (exp((-0.100000 * (V + 32.0))))));
   return t1 + t2;
}
__device__ __forceinline__ double Xr1Mhu(double V) { return 1.0 / (1.0 + (exp(((-26.0 - V) / 7.0)))); }
__device__ __forceinline__ double Xr1TauR(double V)
{
   const double t1 = (1.0 + (exp(((-45.0 - V) / 10.0)))) / 450;
   const double t2 = (1.0 + (exp(((V + 30.0) / 11.5000)))) / 6.0;
   return t1 * t2;
}
__device__ __forceinline__ double Xr2Mhu(double V) { return 1.0 / (1.0 + (exp(((V + 88.0) / 24.0)))); }
__device__ __forceinline__ double Xr2TauR(double V)
{
   const double t1 = (1.0 + (exp(((-60.0 - V) / 20.0)))) / 3.0;
   const double t2 = (1.0 + (exp(((V - 60.0) / 20.0)))) / 1.120;
   return t1 * t2;
}
__device__ __forceinline__ double XsMhu(double V) { return 1.0 / (1.0 + (exp(((-5.0 - V) / 14.0)))); }
__device__ __forceinline__ double XsTauR(double V)
{
   const double t1 = 1400.00 / sqrt(1.0 + exp((5.0 - V) / 6.0));
   const double t2 = 1.0 / (1.0 + (exp(((V - 35.0) / 15.0))));
   return 1.0 / (t1 * t2 + 80.0);
}
__device__ __forceinline__ double rMhu(double V) { return 1.0 / (1.0 + (exp(((20.0 - V) / 6.0)))); }
__device__ __forceinline__ double rTauR(double V)
{
   const double tau = 9.50000 * (exp((-sq((V + 40.0))) / 1800.00)) + 0.800000;
   return 1.0 / tau;
}
__device__ __forceinline__ double dMhu(double V) { return 1.0 / (1.0 + (exp(((-8.0 - V) / 7.50000)))); }
__device__ __forceinline__ double dTauR(double V)
{
   const double t1 = 1.40000 / (1.0 + (exp(((-35.0 - V) / 13.0)))) + 0.250000;
   const double t2 = 1.40000 / (1.0 + (exp(((V + 5.0) / 5.0))));
   const double t3 = 1.0 / (1.0 + (exp(((50.0 - V) / 20.0))));
   return 1 / (t1 * t2 + t3);
}
__device__ __forceinline__ double fMhu(double V) { return 1.0 / (1.0 + (exp(((V + 20.0) / 7.0)))); }
__device__ __forceinline__ double fTauR(double V)
{
   const double tau = 1102.50 * (exp((-sq(V + 27.0) / 225.0))) + 200.0 / (1.0 + (exp(((13.0 - V) / 10.0)))) +
                      180.0 / (1.0 + (exp(((V + 30.0) / 10.0)))) + 20.0;
   return 1 / tau;
}
__device__ __forceinline__ double f2Mhu(double V) { return 0.670000 / (1.0 + (exp(((V + 35.0) / 7.0)))) + 0.330000; }
__device__ __forceinline__ double f2TauR(double V)
{
   const double tau = 562.0 * exp(-sq((V + 27.0)) / 240.0) + 31.0 / (1.0 + (exp(((25.0 - V) / 10.0)))) +
                      80.0 / (1.0 + (exp(((V + 30.0) / 10.0))));
   return 1 / tau;
}
__device__ __forceinline__ double jLMhu(double V) { return exp(-(V + 91.0) * 2 / 6.1); }
__device__ __forceinline__ double jLTauR(double) { return 1.0 / 670.0; }
__device__ __forceinline__ double sMhu0(double V) { return 1.0 / (1.0 + (exp(((V + 28.0) / 5.0)))); }
__device__ __forceinline__ double sMhu1(double V) { return 1.00000 / (1.00000 + (exp(((V + 20.0000) / 5.00000)))); }
__device__ __forceinline__ double sTauR0(double V)
{
   const double tau = 1000.0 * (exp((-sq(V + 67.0) / 1000.0))) + 8.0;
   return 1 / tau;
}
__device__ __forceinline__ double sTauR1(double V)
{
   const double tau = 85.0000 * (exp((-(pow((V + 45.0000), 2.00000)) / 320.000))) +
                      5.00000 / (1.00000 + (exp(((V - 20.0000) / 5.00000)))) + 3.00000;
   return 1 / tau;
}

// TauRecipMod with the host's cubic {x0, x1, c0..c3}: the cubic inside
// [x0, x1], the raw rate outside (elec/TT06Tau.cc:131-151)
__device__ __forceinline__ double smoothed(const double *h, double V, double raw)
{
   if (V < h[0] || V > h[1]) return raw;
   const double x = V - h[0];
   return h[2] + x * (h[3] + x * (h[4] + x * h[5]));
}

}  // namespace exact

// fv0..fv6 exactly (elec/TT06NonGates.c:319-364)
struct ExactVoltageFns {
   const TT06ExactConst *__restrict__ E;
   const TT06Const *__restrict__ K;
   __device__ __forceinline__ void fv05(double V, double fv[6]) const
   {
      const double f1 = E->f1;
      {
         const double expV1 = exp(-f1 * V);
         fv[0] = E->f2 / (1.0 + 0.1245 * exp(-0.1 * V * f1) + 0.0353 * expV1);
      }
      {
         const double expV1 = exp(-f1 * V);
         const double expV5 = exp(E->f5 * V);
         const double expV = expV5 * expV1;
         const double fva = 1.0 / (E->f3 + E->f4 * expV);
         fv[1] = expV5 * fva;
         fv[2] = expV * fva * E->f6;
      }
      {
         const double x = 2.0 * f1 * (V - 15.0);
         const double x2 = x * x;
         fv[3] = x2 < 1e-4 ? (1 + 0.5 * x + x2 / 12.0 - x2 * x2 / 720) * E->f7a
                           : E->f7a * x / (1.0 - exp(-x));
      }
      {
         const double x = 2 * f1 * (V - 15.0);
         const double x2 = x * x;
         fv[4] = x2 < 1e-4 ? (1 - 0.5 * x + x2 / 12.0 - x2 * x2 / 720) * E->f9a
                           : E->f9a * x / (exp(x) - 1.0);
      }
      fv[5] = E->f10 / (1.0 + exp((25.0 - V) / 5.98));
   }
   __device__ __forceinline__ double fv6(double dV0) const
   {
      const double xx = (3.0 * exp(0.0002 * dV0 + 0.02) + exp(0.1 * dV0 - 1.0)) / (1.0 + exp(-0.5 * dV0)) *
                        (10.0 + 10 * exp(0.06 * dV0 - 12.0));
      return K->c12 / (1.0 + xx);
   }
};

// TT06Func::updateGate (elec/TT06Func.cc:64-161) with the exact functions:
// every cell, m by Rush-Larsen, the others by g += dt*(mhu - g)*tauR.
// `gmask` (every bit set) puts each gate in a basic block of its own, as in
// gates64.
__device__ __forceinline__ void gates_exact64(const TT06ExactConst &E, int s_switch, double dt,
                                              double V, double *st, uint32_t gmask)
{
   using namespace exact;
   if (gmask & (1u << 0)) st[S_M] += (mMhu(V) - st[S_M]) * (1 - exp(-dt * mTauR(V)));
   if (gmask & (1u << 1)) {
      const double mhu = hjMhu(V);
      const double raw = hTauR(V);
      const double tauR = E.mod ? smoothed(E.hmod, V, raw) : raw;
      st[S_H] += dt * (mhu - st[S_H]) * tauR;
   }
   if (gmask & (1u << 2)) {
      const double mhu = hjMhu(V);
      const double raw = jTauR(V);
      const double tauR = E.mod ? smoothed(E.jmod, V, raw) : raw;
      st[S_J] += dt * (mhu - st[S_J]) * tauR;
   }
   if (gmask & (1u << 3)) st[S_XR1] += dt * (Xr1Mhu(V) - st[S_XR1]) * Xr1TauR(V);
   if (gmask & (1u << 4)) st[S_XR2] += dt * (Xr2Mhu(V) - st[S_XR2]) * Xr2TauR(V);
   if (gmask & (1u << 5)) st[S_XS] += dt * (XsMhu(V) - st[S_XS]) * XsTauR(V);
   if (gmask & (1u << 6)) st[S_R] += dt * (rMhu(V) - st[S_R]) * rTauR(V);
   if (gmask & (1u << 7)) st[S_D] += dt * (dMhu(V) - st[S_D]) * dTauR(V);
   if (gmask & (1u << 8)) st[S_F] += dt * (fMhu(V) - st[S_F]) * fTauR(V);
   if (gmask & (1u << 9)) st[S_F2] += dt * (f2Mhu(V) - st[S_F2]) * f2TauR(V);
   if (gmask & (1u << 10)) st[S_JL] += dt * (jLMhu(V) - st[S_JL]) * jLTauR(V);
   if (gmask & (1u << 11)) {
      if (s_switch) st[S_S] += dt * (sMhu1(V) - st[S_S]) * sTauR1(V);
      else          st[S_S] += dt * (sMhu0(V) - st[S_S]) * sTauR0(V);
   }
}

}  // namespace cdh
