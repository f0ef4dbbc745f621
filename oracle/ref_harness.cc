/*
 * ref_harness.cc -- drives the reference's own TT06 / FGR kernels, compiled
 * from their sources under the reference tree (see oracle/Makefile, target
 * `ref`).  TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle.
 *
 * It plays the role of TT06Dev_Reaction (elec/TT06Dev_Reaction.cc:83-157,
 * 393-407), which itself does not compile against the live Reaction ABI:
 * build the PADE fit array from (l, m, coef) as scanFit does
 * (TT06Dev_Reaction.cc:557-582), scale the tauR numerators by dt (:130-136),
 * then call update_nonGate and TT06Func::updateGateFast / the v2 kernels.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include "pade.hh"
#include "TT06Func.hh"
#include "TT06Gates.h"
#include "TT06NonGates.h"
#include "FGRUtils.hh"
#include "FibreConductivity.hh"

namespace {
const char *FIT_NAMES[33] = {
   "fv0", "fv1", "fv2", "fv3", "fv4", "fv5", "fv6", "mMhu", "mTauR", "hMhu", "hTauRMod",
   "jMhu", "jTauRMod", "Xr1Mhu", "Xr1TauR", "Xr2Mhu", "Xr2TauR", "XsMhu", "XsTauR",
   "rMhu", "rTauR", "dMhu", "dTauR", "fMhu", "fTauR", "f2Mhu", "f2TauR", "jLMhu",
   "jLTauR", "sMhu0", "sTauR0", "sMhu1", "sTauR1"};
typedef double (*OVF)(double, void *);
/* name -> exact function, as fitFuncMap does (TT06Dev_Reaction.cc:46-59) */
OVF exact_func(int i, int jhTauSmooth)
{
   using namespace TT06Func;
   static const OVF tab[33] = {
      fv0, fv1, fv2, fv3, fv4, fv5, fv6, mMhu, mTauR, hjMhu, hTauRMod, hjMhu, jTauRMod,
      Xr1Mhu, Xr1TauR, Xr2Mhu, Xr2TauR, XsMhu, XsTauR, rMhu, rTauR, dMhu, dTauR, fMhu,
      fTauR, f2Mhu, f2TauR, jLMhu, jLTauR, sMhu0, sTauR0, sMhu1, sTauR1};
   if (!jhTauSmooth && i == 10) return hTauR;
   if (!jhTauSmooth && i == 12) return jTauR;
   return tab[i];
}
PADE g_fit[33];
bool g_init = false;
CellTypeParms g_ct;

void init_once()
{
   if (g_init) return;
   TT06Func::initCnst();
   initExp();
   g_init = true;
}
}

extern "C" {

void ref_init(void) { init_once(); }

double ref_get_c9(void) { init_once(); return get_c9(); }

/* Load a fit set: l[33], m[33], coef packed (sum of l+m per fit).
 * tol <= 0 loads the exact functions instead (configuration A).
 * dt > 0 scales the 13 tauR numerators as the TT06Dev constructor does. */
void ref_load_fits(const int *l, const int *m, const double *coef, double tol,
                   int jhTauSmooth, double dt)
{
   init_once();
   int pos = 0;
   for (int i = 0; i < 33; i++) {
      double V0 = (i == 6) ? 0 : -100.0, V1 = (i == 6) ? 130 : 50.0;
      if (tol > 0) {
         padeApprox(g_fit[i], FIT_NAMES[i], exact_func(i, jhTauSmooth), NULL, 0, 0.1, V0, V1,
                    tol, 0, 0, 0, l[i], m[i], const_cast<double *>(coef + pos));
         pos += l[i] + m[i];
      } else {
         padeApprox(g_fit[i], FIT_NAMES[i], exact_func(i, jhTauSmooth), NULL, 0, 0.1, V0, V1,
                    0.0, 0, 0, 0, 0, 0, NULL);
      }
   }
   if (tol > 0 && dt > 0) {
      PADE *gateFit = g_fit + gateFitOffset;
      for (int eq = 0; eq < 13; eq++) {
         double *tauR = gateFit[2 * eq + 1].coef;
         int mm = gateFit[2 * eq + 1].m;
         for (int j = 0; j < mm; j++) tauR[j] *= dt;
      }
   }
}

double ref_fit_eval(int i, double x) { return g_fit[i].afunc(x, g_fit[i].aparms); }

void ref_set_celltype(int s_switch, double P_NaK, double g_Ks, double g_Kr,
                      double g_to, double g_NaL)
{
   g_ct.cellType = 0;
   g_ct.s_switch = s_switch;
   g_ct.P_NaK = P_NaK; g_ct.g_Ks = g_Ks; g_ct.g_Kr = g_Kr; g_ct.g_to = g_to; g_ct.g_NaL = g_NaL;
   /* midK_i/midNa_i feed only unused locals (TT06NonGates.c:428-433) */
   g_ct.minK_i = g_ct.maxK_i = g_ct.midK_i = 1.0;
   g_ct.minNa_i = g_ct.maxNa_i = g_ct.midNa_i = 1.0;
}

/* state: 19 arrays of n doubles, SoA with row stride `stride` doubles.  The
 * v2 kernels load from 32-byte aligned addresses only, so the caller passes a
 * 32-byte aligned base and a stride that is a multiple of 4, as the
 * reaction's padded buffer guarantees (TT06Dev_Reaction.cc:97, :180-192). */
void ref_nongate(double dt, int n, long stride, double *Vm, double *state, double *dVdt)
{
   double *st[19];
   for (int k = 0; k < 19; k++) st[k] = state + (size_t)k * stride;
   update_nonGate((void *)g_fit, dt, &g_ct, n, Vm, 0, st, dVdt);
}

/* TT06Func::updateGateFast over all 12 gates (configuration B) */
void ref_gates_fast(double dt, int n, long stride, double *Vm, double *state)
{
   double *st[19];
   for (int k = 0; k < 19; k++) st[k] = state + (size_t)k * stride;
   TT06Func::WORK work = {0, n, 0, 12, NULL};
   TT06Func::updateGateFast(dt, n, g_ct.s_switch, Vm, 0, &st[gateOffset], g_fit, work);
}

/* TT06Func::updateGate (generic; with exact functions this is configuration A) */
void ref_gates_generic(double dt, int n, long stride, double *Vm, double *state)
{
   double *st[19];
   for (int k = 0; k < 19; k++) st[k] = state + (size_t)k * stride;
   TT06Func::WORK work = {0, n, 0, 12, NULL};
   TT06Func::updateGate(dt, n, g_ct.s_switch, Vm, 0, &st[gateOffset], g_fit, work);
}

/* one v2 gate kernel, eq in 0..12 (m,h,j,Xr1,Xr2,Xs,r,d,f,f2,jL,s0,s1) */
void ref_gate_v2(double dt, int eq, int n, double *Vm, double *g)
{
   typedef void (*UG)(double, int, double *, double *, double *, double *);
   static const UG k[13] = {update_mGate_v2, update_hGate_v2, update_jGate_v2,
                            update_Xr1Gate_v2, update_Xr2Gate_v2, update_XsGate_v2,
                            update_rGate_v2, update_dGate_v2, update_fGate_v2,
                            update_f2Gate_v2, update_jLGate_v2, update_s0Gate_v2,
                            update_s1Gate_v2};
   PADE *gateFit = g_fit + gateFitOffset;
   k[eq](dt, n, Vm, g, gateFit[2 * eq].coef, gateFit[2 * eq + 1].coef);
}

double ref_exact_func(int i, double V, int jhTauSmooth) { init_once(); return exact_func(i, jhTauSmooth)(V, NULL); }

/* Per-group cell types for ref_react_group (one reaction object per cell
 * type, ReactionManager.cc:184-261). */
static CellTypeParms g_group_ct[8];

void ref_set_group_celltype(int g, int s_switch, double P_NaK, double g_Ks, double g_Kr,
                            double g_to, double g_NaL)
{
   if (g < 0 || g >= 8) return;
   CellTypeParms &c = g_group_ct[g];
   c.cellType = g;
   c.s_switch = s_switch;
   c.P_NaK = P_NaK; c.g_Ks = g_Ks; c.g_Kr = g_Kr; c.g_to = g_to; c.g_NaL = g_NaL;
   c.minK_i = c.maxK_i = c.midK_i = 1.0;
   c.minNa_i = c.maxNa_i = c.midNa_i = 1.0;
}

/* TT06Dev_Reaction::calc (TT06Dev_Reaction.cc:393-407) in configuration B on
 * one reaction group's SoA state: update_nonGate, then updateGateFast (the
 * v2 kernels).  The signature is the oracle loop's reaction hook
 * (ora_react_fn), so the loop can time the reference's own kernels. */
void ref_react_group(void *ctx, int group, double dt, int n, double *Vm, double *state,
                     int64_t stride, double *dVdt)
{
   (void)ctx;
   double *st[19];
   for (int k = 0; k < 19; k++) st[k] = state + (size_t)k * stride;
   CellTypeParms *ct = &g_group_ct[group];
   update_nonGate((void *)g_fit, dt, ct, n, Vm, 0, st, dVdt);
   TT06Func::WORK work = {0, n, 0, 12, NULL};
   TT06Func::updateGateFast(dt, n, ct->s_switch, Vm, 0, &st[gateOffset], g_fit, work);
}

/* One row of TT06Func::getStandardCellTypes (TT06Func.cc:421-612):
 * parms = {P_NaK, g_Ks, g_Kr, g_to, g_NaL, Vm}, state[19] by each STATE's
 * index (TT06_DEV_STATE_INDEX, TT06Func.h:10-14; K_i itself in slot 1).
 * Returns the number of states the row lists, or -1 for an unknown name. */
int ref_celltype_standard(const char *name, int *s_switch, double *parms, double *state)
{
   init_once();
   std::map<std::string, TT06Func::CellTypeParmsFull> t = TT06Func::getStandardCellTypes();
   auto it = t.find(name);
   if (it == t.end()) return -1;
   const TT06Func::CellTypeParmsFull &c = it->second;
   *s_switch = c.s_switch;
   parms[0] = c.P_NaK; parms[1] = c.g_Ks; parms[2] = c.g_Kr;
   parms[3] = c.g_to; parms[4] = c.g_NaL; parms[5] = c.Vm;
   int cnt = 0;
   for (const auto &kv : c.state) {
      if (kv.second.index < 0 || kv.second.index >= 19) return -1;
      state[kv.second.index] = kv.second.value;
      cnt++;
   }
   return cnt;
}

/* names getStandardCellTypes defines, space separated, into buf */
int ref_celltype_names(char *buf, int len)
{
   init_once();
   std::string s;
   for (const auto &kv : TT06Func::getStandardCellTypes()) s += kv.first + " ";
   if ((int)s.size() + 1 > len) return -1;
   std::memcpy(buf, s.c_str(), s.size() + 1);
   return (int)TT06Func::getStandardCellTypes().size();
}

void ref_fgr_f2(unsigned iFace, int *tissue, double *gradPhi /* 3*19 */)
{
   double g[3][19];
   std::memset(g, 0, sizeof(g));
   FGRUtils::f2(iFace, tissue, g);
   std::memcpy(gradPhi, g, sizeof(g));
}

void ref_fibre_conductivity(double theta, double phi, double sigmaL, double sigmaT,
                            double *out /* a11 a12 a13 a22 a23 a33 */)
{
   FibreConductivityParms p;
   p.sigmaLi = sigmaL;
   p.sigmaTi = sigmaT;
   FibreConductivity fc(p);
   SymmetricTensor s;
   fc.compute(theta, phi, s);
   out[0] = s.a11; out[1] = s.a12; out[2] = s.a13;
   out[3] = s.a22; out[4] = s.a23; out[5] = s.a33;
}

} /* extern "C" */
