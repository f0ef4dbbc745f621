/*
 * tt06_oracle.c -- CPU restatement of the TT06Dev reaction (TEST INFRASTRUCTURE ONLY).
 *
 * Arithmetic is kept in the reference's evaluation order, operation by
 * operation, and this file is compiled with -ffp-contract=off so that no
 * multiply-add is fused (the reference's x86 build does not fuse either).
 * That is what makes a bitwise comparison with oracle/_ref possible.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Constants: elec/TT06NonGates.c:11-251.  The evaluation order of every     */
/* derived constant follows the reference so the doubles are identical.     */
/* ------------------------------------------------------------------------ */
typedef struct {
   double c2, c3, c4, c5, c6, c7, c8, c9, c11, c12, c13, c14, c15, c16, c17,
          c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29, c30,
          c31, c32, c33, c34, c36, c40, c43, c44;
   double f1, f2, f3, f4, f5, f6, f7, f7a, f9, f9a, f10;
} tt06_const;

static tt06_const K;
static double exp_coef[17];     /* 1/i!, elec/TT06GatesSimd.c:10-14 */
static int inited = 0;

#define CUBE3(x) ((x) * (x) * (x))

void ora_tt06_init(void)
{
   if (inited) return;
   /* physical parameters, TT06 (2006) CellML values */
   const double R = 8314.472, T = 310, F = 96485.3415, Cm = 0.185, V_c = 0.016404;
   const double P_kna = 0.03, K_o = 5.4, Na_o = 140, Ca_o = 2, g_K1 = 5.405;
   const double g_bca = 0.000592, g_CaL = 0.0000398, g_pCa = 0.1238;
   const double K_mk = 1, K_mNa = 40, K_NaCa = 1000, K_sat = 0.1, alpha = 2.5;
   const double gamma_ = 0.35, Km_Ca = 1.38, Km_Nai = 87.5, K_pCa = 0.0005, g_pK = 0.0146;
   const double k1p = 0.15, k2p = 0.045, k3 = 0.06, k4 = 0.005, EC = 1.5;
   const double max_sr = 2.5, min_sr = 1, V_rel = 0.102, V_xfer = 0.0038, K_up = 0.00025;
   const double V_leak = 0.00036, Vmax_up = 0.006375, Buf_c = 0.2, K_buf_c = 0.001;
   const double Buf_sr = 10, K_buf_sr = 0.3, Buf_ss = 0.4, K_buf_ss = 0.00025;
   const double V_sr = 0.001094, V_ss = 0.00005468;
   const double g_Na = 14.838, g_bna = 0.00029;

   double c1 = F / (R * T);
   K.c2 = P_kna;
   K.c3 = -1 / c1;
   K.c4 = -K.c3 * log(Na_o);
   K.c5 = -K.c3 * log(K_o);
   K.c6 = -K.c3 * log(K_o + P_kna * Na_o);
   K.c8 = -0.5 * K.c3 * log(Ca_o);
   K.c9 = -Cm / (V_c * F);
   K.c7 = g_bca;
   K.c11 = sqrt(K_o / 5.4);
   K.c12 = g_K1 * sqrt(K_o / 5.4);
   K.c13 = Cm / (2.0 * V_ss * F * K.c9);
   K.c14 = V_sr / V_ss;
   K.c15 = -V_ss / V_c;
   K.c16 = V_sr / V_c;
   K.c17 = k3 / (k1p * k2p) * K.c9;
   K.c18 = k2p * max_sr / K.c9;
   K.c19 = -k2p * (max_sr - min_sr) / K.c9;
   K.c20 = g_Na;
   K.c21 = g_bna;
   K.c22 = 1 / K.c9;
   K.c23 = V_xfer / (K.c15 * K.c9);
   K.c24 = g_pCa;
   K.c25 = 1.0 / K_mNa;
   K.c26 = 1.0 / K_pCa;
   K.c27 = 1.0 / K_up;
   K.c28 = 1.0 / sqrt(Buf_c * K_buf_c);
   K.c29 = K_buf_c * K.c28;
   K.c30 = 1.0 / EC;
   K.c31 = 1.0 / sqrt(Buf_sr * K_buf_sr);
   K.c32 = K_buf_sr * K.c31;
   K.c33 = 1.0 / sqrt(Buf_ss * K_buf_ss);
   K.c34 = K_buf_ss * K.c33;
   K.c36 = k4 / K.c9;
   K.c40 = V_rel / K.c9;
   K.c43 = V_leak / K.c9;
   K.c44 = Vmax_up / K.c9;

   K.f1 = c1;
   K.f2 = -2.0 * K_o / (K_o + K_mk);
   K.f3 = ((CUBE3(Km_Nai) + CUBE3(Na_o)) * (Km_Ca + Ca_o)) / (K_NaCa * Ca_o);
   K.f4 = K.f3 * K_sat;
   K.f5 = gamma_ * K.f1;
   K.f6 = (CUBE3(Na_o) * alpha / Ca_o);
   K.f7 = g_CaL * F * K.f1;
   K.f7a = 0.5 * g_CaL * F;
   K.f9 = 4.0 * Ca_o;
   K.f9a = 4.0 * Ca_o * K.f7a;
   K.f10 = g_pK;

   exp_coef[0] = 1.0;
   for (int i = 1; i < 17; i++) exp_coef[i] = exp_coef[i - 1] / i;
   inited = 1;
}

double ora_c9(void) { ora_tt06_init(); return K.c9; }

/* ------------------------------------------------------------------------ */
/* Cell types: elec/TT06Func.cc:421-612                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
   const char *name;
   int s_switch;
   double P_NaK, g_Ks, g_Kr, g_to, g_NaL, Vm;
   /* Ca_i, K_i, Na_i, Ca_ss, Ca_SR, R', fCass, m, h, j, Xr1, Xr2, Xs, r, d, f, f2, jL, s */
   double s[ORA_NSTATE];
} celltype_row;

static const celltype_row CELLTYPES[] = {
   {"endoCellML", 0, 2.724, 0.392, 0.153, 0.073, 0.0, -86.709,
    {0.00013, 138.4, 10.355, 0.00036, 3.715, 0.9068, 0.9953, 0.00155, 0.7573, 0.7225,
     0.00448, 0.476, 0.0087, 2.235e-8, 3.164e-5, 0.8009, 0.9778, 0.066, 0.3212}},
   {"midCellML", 1, 2.724, 0.098, 0.153, 0.294, 0.0, -85.423,
    {0.000153, 138.52, 10.132, 0.00042, 4.272, 0.8978, 0.9942, 0.00165, 0.749, 0.6788,
     0.0165, 0.473, 0.0174, 2.347e-8, 3.288e-5, 0.7026, 0.9526, 0.066, 0.999998}},
   {"epiCellML", 1, 2.724, 0.392, 0.153, 0.294, 0.0, -85.23,
    {0.000126, 136.89, 8.604, 0.00036, 3.64, 0.9073, 0.9953, 0.00172, 0.7444, 0.7045,
     0.00621, 0.4712, 0.0095, 2.42e-8, 3.373e-5, 0.7888, 0.9755, 0.066, 0.999998}},
   {"endoRRG", 0, 3.000, 0.392, 0.153, 0.073, 0.15, -86.709,
    {0.00013, 138.4, 10.355, 0.00036, 3.715, 0.9068, 0.9953, 0.00155, 0.7573, 0.7225,
     0.00448, 0.476, 0.0087, 2.235e-8, 3.164e-5, 0.8009, 0.9778, 0.066, 0.3212}},
   {"midRRG", 1, 3.100, 0.098, 0.153, 0.294, 0.3, -85.423,
    {0.000153, 138.52, 10.132, 0.00042, 4.272, 0.8978, 0.9942, 0.00165, 0.749, 0.6788,
     0.0165, 0.473, 0.0174, 2.347e-8, 3.288e-5, 0.7026, 0.9526, 0.066, 0.999998}},
   {"epiRRG", 1, 3.000, 0.392, 0.153, 0.294, 0.15, -85.23,
    {0.000126, 136.89, 8.604, 0.00036, 3.64, 0.9073, 0.9953, 0.00172, 0.7444, 0.7045,
     0.00621, 0.4712, 0.0095, 2.42e-8, 3.373e-5, 0.7888, 0.9755, 0.066, 0.999998}},
};

int ora_celltype_standard(const char *name, ora_celltype *out)
{
   for (unsigned k = 0; k < sizeof(CELLTYPES) / sizeof(CELLTYPES[0]); k++) {
      if (strcmp(CELLTYPES[k].name, name) != 0) continue;
      const celltype_row *r = &CELLTYPES[k];
      out->s_switch = r->s_switch;
      out->P_NaK = r->P_NaK; out->g_Ks = r->g_Ks; out->g_Kr = r->g_Kr;
      out->g_to = r->g_to; out->g_NaL = r->g_NaL; out->Vm0 = r->Vm;
      memcpy(out->state0, r->s, sizeof(r->s));
      return 0;
   }
   return -1;
}

/* state init: elec/TT06Dev_Reaction.cc:194-199 */
void ora_tt06_init_state(const ora_celltype *ct, int n, double *state,
                         int64_t stride, double *Vm)
{
   double c9 = ora_c9();
   for (int i = 0; i < n; i++) {
      for (int s = 0; s < ORA_NSTATE; s++) state[s * stride + i] = ct->state0[s];
      state[ORA_K_I * stride + i] = ct->state0[ORA_K_I] / c9 + ct->Vm0;
      if (Vm) Vm[i] = ct->Vm0;
   }
}

/* ------------------------------------------------------------------------ */
/* Rational fits: elec/pade.cc:17-36                                         */
/* ------------------------------------------------------------------------ */
static inline double horner(const double *a, int lo, int hi, double x)
{
   /* a[hi] + x*(...) down to a[lo]; the first step a[hi] + x*0 is a[hi] */
   double s = a[hi];
   for (int j = hi - 1; j >= lo; j--) s = a[j] + x * s;
   return s;
}

double ora_pade_eval(const ora_pade *p, double x)
{
   double num = horner(p->coef, 0, p->m - 1, x);
   if (p->l < 2) return num;                      /* polyFunc */
   double den = horner(p->coef, p->m, p->m + p->l - 1, x);
   return num / den;
}

/* elec/TT06Dev_Reaction.cc:130-136: the tauR numerators absorb dt */
void ora_fitset_scale_dt(ora_fitset *fs, double dt)
{
   if (fs->dt_scaled) return;
   for (int eq = 0; eq < 13; eq++) {
      ora_pade *p = &fs->fit[7 + 2 * eq + 1];
      for (int j = 0; j < p->m; j++) p->coef[j] *= dt;
   }
   fs->dt_scaled = 1;
}

/* ------------------------------------------------------------------------ */
/* Non-gate update, elec/TT06NonGates.c:374-563 (generic fv05/fv6 path)      */
/* ------------------------------------------------------------------------ */
typedef void (*fvfun)(const void *ctx, double Vm, double fv[6]);
typedef double (*fv6fun)(const void *ctx, double dV0);

static void fv05_fit(const void *ctx, double Vm, double fv[6])
{
   const ora_fitset *fs = (const ora_fitset *)ctx;
   for (int k = 0; k < 6; k++) fv[k] = ora_pade_eval(&fs->fit[k], Vm);
}
static double fv6_fit(const void *ctx, double dV0)
{
   const ora_fitset *fs = (const ora_fitset *)ctx;
   return ora_pade_eval(&fs->fit[6], dV0);
}

/* exact voltage functions, elec/TT06NonGates.c:319-364 */
static double fv0_x(double V) { double e1 = exp(-K.f1 * V); return K.f2 / (1.0 + 0.1245 * exp(-0.1 * V * K.f1) + 0.0353 * e1); }
static double fv1_x(double V) { double e1 = exp(-K.f1 * V), e5 = exp(K.f5 * V), e = e5 * e1; double a = 1.0 / (K.f3 + K.f4 * e); return e5 * a; }
static double fv2_x(double V) { double e1 = exp(-K.f1 * V), e5 = exp(K.f5 * V), e = e5 * e1; double a = 1.0 / (K.f3 + K.f4 * e); return e * a * K.f6; }
static double fv3_x(double V)
{
   double x = 2.0 * K.f1 * (V - 15.0), x2 = x * x;
   if (x2 < 1e-4) return (1 + 0.5 * x + x2 / 12.0 - x2 * x2 / 720) * K.f7a;
   return K.f7a * x / (1.0 - exp(-x));
}
static double fv4_x(double V)
{
   double x = 2 * K.f1 * (V - 15.0), x2 = x * x;
   if (x2 < 1e-4) return (1 - 0.5 * x + x2 / 12.0 - x2 * x2 / 720) * K.f9a;
   return K.f9a * x / (exp(x) - 1.0);
}
static double fv5_x(double V) { return K.f10 / (1.0 + exp((25.0 - V) / 5.98)); }
static double fv6_x(double d)
{
   double xx = (3.0 * exp(0.0002 * d + 0.02) + exp(0.1 * d - 1.0)) / (1.0 + exp(-0.5 * d)) * (10.0 + 10 * exp(0.06 * d - 12.0));
   return K.c12 / (1.0 + xx);
}
static void fv05_exact(const void *ctx, double V, double fv[6])
{
   (void)ctx;
   fv[0] = fv0_x(V); fv[1] = fv1_x(V); fv[2] = fv2_x(V);
   fv[3] = fv3_x(V); fv[4] = fv4_x(V); fv[5] = fv5_x(V);
}
static double fv6_exact(const void *ctx, double d) { (void)ctx; return fv6_x(d); }

#define SIGM(x) ((x) / (1.0 + (x)))

static void nongate_core(fvfun fv05, fv6fun fv6f, const void *ctx, double dt,
                         const ora_celltype *ct, const ora_scales *sc, int n,
                         const double *VM, double *st, int64_t S, double *dVdt)
{
   ora_tt06_init();
   const double s_kr = sc ? sc->kr : 1.0, s_cal = sc ? sc->cal : 1.0, s_na = sc ? sc->na : 1.0;
   /* current scales, elec/TT06NonGates.c:405-426 (the literal 1s kept as 1.0) */
   const double c_K1 = 1, c_Na = s_na * K.c20, c_bNa = 1 * K.c21, c_CaL = s_cal;
   const double c_bCa = 1 * K.c7, c_NaCa = 1, c_pCa = 1 * K.c24, c_pK = 1;
   const double c_NaK = 1 * ct->P_NaK, c_Ks = 1 * ct->g_Ks, c_Kr = s_kr * ct->g_Kr;
   const double c_to = 1 * ct->g_to, c_NaL = 1 * ct->g_NaL;

   for (int ii = 0; ii < n; ii++) {
      const double Vm = VM[ii];
      double fv[6];
      fv05(ctx, Vm, fv);
      const double Na_i = st[ORA_NA_I * S + ii], Ca_i = st[ORA_CA_I * S + ii];
      const double Ca_ss = st[ORA_CA_SS * S + ii], Ca_SR = st[ORA_CA_SR * S + ii];
      const double fCass = st[ORA_FCASS * S + ii], R_prime = st[ORA_R_PRIME * S + ii];
      double dVK_i = st[ORA_K_I * S + ii];
      const double m = st[ORA_M * S + ii], h = st[ORA_H * S + ii], j = st[ORA_J * S + ii];
      const double xr1 = st[ORA_XR1 * S + ii], xr2 = st[ORA_XR2 * S + ii], xs = st[ORA_XS * S + ii];
      const double r = st[ORA_R * S + ii], d = st[ORA_D * S + ii], f = st[ORA_F * S + ii];
      const double f2 = st[ORA_F2 * S + ii], jL = st[ORA_JL * S + ii], s = st[ORA_S * S + ii];

      double dVR = 0.0;
      /* calcium block */
      const double x1 = Ca_i * K.c26;
      const double x2 = (Ca_i * K.c27) * (Ca_i * K.c27);
      const double x3 = (Ca_i * K.c28 + K.c29) * (Ca_i * K.c28 + K.c29);
      const double sg1 = SIGM(x1), sg2 = SIGM(x2), sg3 = SIGM(x3);
      const double dV3 = Vm - 0.5 * K.c3 * log(Ca_i) - K.c8;
      const double I_NaCa = c_NaCa * (CUBE3(Na_i) * fv[1] - Ca_i * fv[2]);
      const double I_bCa = c_bCa * dV3;
      const double I_pCa = c_pCa * sg1;
      const double I_up = K.c44 * sg2;
      const double I_leak = K.c43 * (Ca_SR - Ca_i);
      const double I_xfer = K.c23 * (Ca_ss - Ca_i);
      const double I_delta = I_leak - I_up;
      const double I_sum = I_bCa + I_pCa;
      st[ORA_CA_I * S + ii] = Ca_i + (dt * K.c9) * (sg3 * (0.5 * I_sum - I_NaCa + I_xfer * K.c15 + I_delta * K.c16));
      dVR -= I_sum;

      /* potassium / sodium block */
      const double K_i = K.c9 * (dVK_i - Vm);
      const double x0 = Na_i * K.c25;
      const double dV0 = Vm - K.c3 * log(K_i) - K.c5;
      const double dV1 = Vm - K.c3 * log(Na_i) - K.c4;
      const double dV2 = dV0 - K.c3 * log(1 + K.c2 * Na_i / K_i) + K.c5 - K.c6;
      const double fv6v = fv6f(ctx, dV0);
      const double I_NaK = c_NaK * SIGM(x0) * fv[0];
      const double I_pK = c_pK * fv[5] * dV0;
      const double I_K1 = c_K1 * fv6v * dV0;
      const double I_to = c_to * r * s * dV0;
      const double I_Kr = c_Kr * xr1 * xr2 * dV0;
      const double I_Na = c_Na * CUBE3(m) * h * j * dV1;
      const double I_NaL = c_NaL * CUBE3(m) * jL * dV1;
      const double I_bNa = c_bNa * dV1;
      const double I_Ks = c_Ks * (xs * xs) * dV2;
      const double iNa = 3 * I_NaCa - 1.5 * I_NaK + I_Na + I_bNa + I_NaL;
      const double iK = I_Ks + I_NaK + I_pK + I_K1 + I_to + I_Kr;
      dVK_i += dt * iK;
      dVR -= iNa + iK - 2.0 * I_NaCa;
      st[ORA_NA_I * S + ii] = Na_i + (dt * K.c9) * iNa;

      /* Ca_ss, Ca_SR, R', fCass block */
      const double x4 = (Ca_SR * K.c30) * (Ca_SR * K.c30);
      const double x5 = (Ca_SR * K.c31 + K.c32) * (Ca_SR * K.c31 + K.c32);
      const double x6 = (Ca_ss * K.c33 + K.c34) * (Ca_ss * K.c33 + K.c34);
      const double sg4 = SIGM(x4), sg5 = SIGM(x5), sg6 = SIGM(x6);
      const double tmp8 = (K.c18 + K.c19 * sg4);
      const double tmp9 = tmp8 * Ca_ss + K.c36;
      const double I_CaL = c_CaL * d * f * f2 * fCass * (fv[4] - Ca_ss * fv[3]);
      const double O = (Ca_ss * Ca_ss) * R_prime / (tmp8 * K.c17 + (Ca_ss * Ca_ss));
      const double I_rel = K.c40 * O * (Ca_SR - Ca_ss);
      st[ORA_CA_SS * S + ii] = Ca_ss + (dt * K.c9) * sg6 * (I_xfer + I_rel * K.c14 + I_CaL * K.c13);
      st[ORA_CA_SR * S + ii] = Ca_SR - (dt * K.c9) * sg5 * (I_delta + I_rel);
      st[ORA_R_PRIME * S + ii] = R_prime + (dt * K.c9) * (K.c36 - tmp9 * R_prime);
      const double t1 = 1.0 / (1.0 + (20 * Ca_ss) * (20 * Ca_ss));
      const double mhu = 0.600000 * t1 + 0.4000000;
      const double tauR = 1.0 / (80.0 * t1 + 2.0);
      st[ORA_FCASS * S + ii] = fCass + dt * (mhu - fCass) * tauR;
      dVR += I_CaL;

      st[ORA_K_I * S + ii] = dVK_i + dt * dVR;
      dVdt[ii] = dVR;
   }
}

void ora_tt06_nongate(const ora_fitset *fs, double dt, const ora_celltype *ct,
                      const ora_scales *sc, int n, const double *Vm,
                      double *state, int64_t stride, double *dVdt)
{
   nongate_core(fv05_fit, fv6_fit, fs, dt, ct, sc, n, Vm, state, stride, dVdt);
}

void ora_tt06_nongate_exact(double dt, const ora_celltype *ct, const ora_scales *sc,
                            int n, const double *Vm, double *state, int64_t stride,
                            double *dVdt)
{
   nongate_core(fv05_exact, fv6_exact, NULL, dt, ct, sc, n, Vm, state, stride, dVdt);
}

/* ------------------------------------------------------------------------ */
/* Gate updates, configuration B ("v2" kernels, elec/TT06GatesSimd.c).       */
/*                                                                           */
/* Every v2 kernel evaluates mhu = num/den and tauRdt = dt*num'/den' by      */
/* Horner in the same order as pade.cc, then applies one of three update     */
/* forms (written here as the kernel writes them with unfused mul/add):     */
/*   RL  (m gate, :239-241, 468-484):   e = sum_{i<=16} x^i/i!, r = 1 - 1/e, */
/*                                       g = r*(mhu - g) + g                 */
/*   DIF (h,j,Xr1,Xs,r,d,f,f2,jL,s0,s1): g = r*mhu + (g - r*g)               */
/*   SUM (Xr2, :1645-1651):              g = (mhu*r + g) - g*r               */
/*   SUB (f gate, lanes 0-3 of each 12-cell group of its main loop,         */
/*        :2736-2741):                   g = (mhu - g)*r + g                 */
/* The v2 loops advance 4 cells at a time and stop at the last full group   */
/* (TTROUND, :7): cells past 4*floor(n/4) keep their gates.  The f-gate     */
/* main loop covers cells below 12*floor(n/12); its 4-wide tail loop uses   */
/* DIF.  So the form of a cell's f-gate update depends on its index i in    */
/* its reaction group: see ora_gate_flags().                                */
/* ------------------------------------------------------------------------ */
enum { UPD_RL = 0, UPD_DIF = 1, UPD_SUM = 2, UPD_SUB = 3 };
static const int GATE_FORM[13] = {
   UPD_RL,  /* m   */ UPD_DIF, /* h  */ UPD_DIF, /* j  */ UPD_DIF, /* Xr1 */
   UPD_SUM, /* Xr2 */ UPD_DIF, /* Xs */ UPD_DIF, /* r  */ UPD_DIF, /* d   */
   UPD_DIF, /* f   */ UPD_DIF, /* f2 */ UPD_DIF, /* jL */ UPD_DIF, /* s0  */
   UPD_DIF  /* s1  */
};

static inline double gate_step(int form, double g, double mhu, double x)
{
   if (form == UPD_RL) {
      double e = horner(exp_coef, 0, 16, x);
      double r = 1.0 - 1.0 / e;
      return r * (mhu - g) + g;
   }
   if (form == UPD_SUM) return (mhu * x + g) - g * x;
   if (form == UPD_SUB) return (mhu - g) * x + g;
   return x * mhu + (g - x * g);
}

/* Per-cell flags from the cell's index i in a reaction group of n cells:
 *   bit 0: the v2 loops never reach the cell, its gates stay fixed;
 *   bit 1: its f gate takes the SUB form. */
int ora_gate_flags(int64_t i, int64_t n)
{
   int f = 0;
   if (i >= 4 * (n / 4)) f |= 1;
   if (i < 12 * (n / 12) && (i % 12) < 4) f |= 2;
   return f;
}

static void gate_v2_core(const ora_fitset *fs, int eq, int n, const double *Vm, double *g,
                         int loop_rule)
{
   ora_tt06_init();
   const ora_pade *pm = &fs->fit[7 + 2 * eq], *pt = &fs->fit[7 + 2 * eq + 1];
   for (int i = 0; i < n; i++) {
      const int fl = loop_rule ? ora_gate_flags(i, n) : 0;
      if (fl & 1) continue;
      const int form = (eq == 8 && (fl & 2)) ? UPD_SUB : GATE_FORM[eq];
      double mhu = ora_pade_eval(pm, Vm[i]);
      double x = ora_pade_eval(pt, Vm[i]);
      g[i] = gate_step(form, g[i], mhu, x);
   }
}

void ora_tt06_gate_v2(const ora_fitset *fs, int eq, int n, const double *Vm, double *g)
{
   gate_v2_core(fs, eq, n, Vm, g, 1);
}

void ora_tt06_gates_v2(const ora_fitset *fs, int s_switch, int n,
                       const double *Vm, double *state, int64_t stride)
{
   for (int eq = 0; eq < 11; eq++)
      ora_tt06_gate_v2(fs, eq, n, Vm, state + (ORA_M + eq) * stride);
   ora_tt06_gate_v2(fs, s_switch ? 12 : 11, n, Vm, state + ORA_S * stride);
}

/* The same update forms with every cell of the group advanced and the f gate
 * in the tail loop's form: the gates of cells whose position in their
 * reaction group depends on how the anatomy is cut (no counterpart in the
 * reference, whose groups are per rank). */
void ora_tt06_gates_v2_all(const ora_fitset *fs, int s_switch, int n,
                           const double *Vm, double *state, int64_t stride)
{
   for (int eq = 0; eq < 11; eq++)
      gate_v2_core(fs, eq, n, Vm, state + (ORA_M + eq) * stride, 0);
   gate_v2_core(fs, s_switch ? 12 : 11, n, Vm, state + ORA_S * stride, 0);
}

/* ------------------------------------------------------------------------ */
/* Configuration A: exact gate functions, elec/TT06Func.cc:64-161, 261-417   */
/* ------------------------------------------------------------------------ */
static double Xr1Mhu(double V) { return 1.0 / (1.0 + (exp(((-26.0 - V) / 7.0)))); }
static double Xr1TauR(double V)
{
   double t1 = (1.0 + (exp(((-45.0 - V) / 10.0)))) / 450;
   double t2 = (1.0 + (exp(((V + 30.0) / 11.5000)))) / 6.0;
   return t1 * t2;
}
static double Xr2Mhu(double V) { return 1.0 / (1.0 + (exp(((V + 88.0) / 24.0)))); }
static double Xr2TauR(double V)
{
   double t1 = (1.0 + (exp(((-60.0 - V) / 20.0)))) / 3.0;
   double t2 = (1.0 + (exp(((V - 60.0) / 20.0)))) / 1.120;
   return t1 * t2;
}
static double XsMhu(double V) { return 1.0 / (1.0 + (exp(((-5.0 - V) / 14.0)))); }
static double XsTauR(double V)
{
   double t1 = 1400.00 / sqrt(1.0 + exp((5.0 - V) / 6.0));
   double t2 = 1.0 / (1.0 + (exp(((V - 35.0) / 15.0))));
   return 1.0 / (t1 * t2 + 80.0);
}
static double mMhu(double V) { double a = 1.0 + exp((-56.8600 - V) / 9.03000); return 1.0 / (a * a); }
static double mTauR(double V)
{
   double t1 = 1.0 / (1.0 + (exp(((-60.0 - V) / 5.0))));
   double t2 = 0.10000 / (1.0 + (exp(((V + 35.0) / 5.0)))) + 0.100000 / (1.0 + (exp(((V - 50.0) / 200.0))));
   return 1.0 / (t1 * t2);
}
static double hjMhu(double V) { double a = (1.0 + (exp(((V + 71.5500) / 7.43000)))); return 1.0 / (a * a); }
static double hTauR(double V)
{
   double t1 = (V < -40.0 ? 0.0570000 * (exp((-(V + 80.0) / 6.80000))) : 0.0);
   double t2 = (V < -40.0 ? 2.70000 * (exp((0.0790000 * V))) + 310000. * (exp((0.348500 * V)))
                          : 0.770000 / (0.130000 * (1.0 + (exp(((V + 10.6600) / -11.1000))))));
   return t1 + t2;
}
static double jTauR(double V)
{
   double t1 = (V < -40.0 ? (((-25428.0 * (exp((0.244400 * V))) - 6.94800e-06 * (exp((-0.0439100 * V)))) * (V + 37.7800)) / 1.0) / (1.0 + (exp((0.311000 * (V + 79.2300))))) : 0.0);
   double t2 = (V < -40.0 ? (0.0242400 * (exp((-0.0105200 * V)))) / (1.0 + (exp((-0.137800 * (V + 40.1400)))))
                          : (0.600000 * (exp((0.0570000 * V)))) / (1.0 + (exp((-0.100000 * (V + 32.0))))));
   return t1 + t2;
}
static double rMhu(double V) { return 1.0 / (1.0 + (exp(((20.0 - V) / 6.0)))); }
static double rTauR(double V) { double tau = 9.50000 * (exp((-((V + 40.0) * (V + 40.0))) / 1800.00)) + 0.800000; return 1.0 / tau; }
static double dMhu(double V) { return 1.0 / (1.0 + (exp(((-8.0 - V) / 7.50000)))); }
static double dTauR(double V)
{
   double t1 = 1.40000 / (1.0 + (exp(((-35.0 - V) / 13.0)))) + 0.250000;
   double t2 = 1.40000 / (1.0 + (exp(((V + 5.0) / 5.0))));
   double t3 = 1.0 / (1.0 + (exp(((50.0 - V) / 20.0))));
   return 1 / (t1 * t2 + t3);
}
static double fMhu(double V) { return 1.0 / (1.0 + (exp(((V + 20.0) / 7.0)))); }
static double fTauR(double V)
{
   double tau = 1102.50 * (exp((-((V + 27.0) * (V + 27.0)) / 225.0))) + 200.0 / (1.0 + (exp(((13.0 - V) / 10.0)))) + 180.0 / (1.0 + (exp(((V + 30.0) / 10.0)))) + 20.0;
   return 1 / tau;
}
static double f2Mhu(double V) { return 0.670000 / (1.0 + (exp(((V + 35.0) / 7.0)))) + 0.330000; }
static double f2TauR(double V)
{
   double tau = 562.0 * exp(-((V + 27.0) * (V + 27.0)) / 240.0) + 31.0 / (1.0 + (exp(((25.0 - V) / 10.0)))) + 80.0 / (1.0 + (exp(((V + 30.0) / 10.0))));
   return 1 / tau;
}
static double jLMhu(double V) { return exp(-(V + 91.0) * 2 / 6.1); }
static double jLTauR(double V) { (void)V; return 1.0 / 670.0; }
static double sMhu0(double V) { return 1.0 / (1.0 + (exp(((V + 28.0) / 5.0)))); }
static double sMhu1(double V) { return 1.00000 / (1.00000 + (exp(((V + 20.0000) / 5.00000)))); }
static double sTauR0(double V) { double tau = 1000.0 * (exp((-((V + 67.0) * (V + 67.0)) / 1000.0))) + 8.0; return 1 / tau; }
static double sTauR1(double V)
{
   double tau = 85.0000 * (exp((-(pow((V + 45.0000), 2.00000)) / 320.000))) + 5.00000 / (1.00000 + (exp(((V - 20.0000) / 5.00000)))) + 3.00000;
   return 1 / tau;
}

/* Smoothed h and j rates (mod = 1, jhTauSmooth): TauRecipMod
 * (elec/TT06Tau.cc:131-151) replaces 1/tau on [x0, x1] by the cubic Hermite
 * interpolant of 1/tau and its derivative at x0 and x1, the two branches
 * of hTauRecip / jTauRecip (:14-129).  The derivatives are needed only at
 * the two ends, where they set the cubic; their operation order follows the
 * reference, so the cubic's coefficients are the same doubles. */
static double h_recip(double V, double *dR)
{
   double t1, t2, dt1, dt2;
   if (V < -40.0) {
      t1 = 0.0570000 * exp(-(V + 80.0) / 6.80000);
      dt1 = -t1 / 6.8000;
      t2 = 2.70000 * exp(0.0790000 * V) + 310000. * exp(0.348500 * V);
      dt2 = 0.079 * 2.70000 * exp(0.0790000 * V) + 0.3485 * 310000. * exp(0.348500 * V);
   } else {
      const double c = (0.770000 / 0.130000);
      const double e = exp((V + 10.6600) / -11.1000);
      const double s = 1.0 / (1.0 + e);
      t1 = 0.0;
      dt1 = 0.0;
      t2 = c * s;
      dt2 = c * s * s * e / 11.1;
   }
   *dR = dt1 + dt2;
   return t1 + t2;
}
static double j_recip(double V, double *dR)
{
   double t1, t2, dt1, dt2;
   if (V < -40.0) {
      double a1 = (-25428.0 * exp(0.244400 * V) - 6.94800e-06 * exp(-0.0439100 * V));
      double da1 = (-0.24400 * 25428.0 * exp(0.244400 * V) + 0.0439100 * 6.94800e-06 * exp(-0.0439100 * V));
      double a2 = (V + 37.7800);
      double da2 = 1.0;
      double c = 0.311;
      double e = exp(c * (V + 79.2300));
      double a3 = 1 / (1.0 + e);
      double da3 = -a3 * a3 * e * c;
      t1 = a1 * a2 * a3;
      double q = da1 / a1 + da2 / a2 + da3 / a3;
      dt1 = q * t1;
      a1 = 0.0242400 * exp(-0.0105200 * V);
      da1 = -0.0105200 * a1;
      c = -0.137800;
      e = exp(c * (V + 40.1400));
      a2 = 1.0 / (1.0 + e);
      da2 = -c * a2 * a2 * e;
      t2 = a1 * a2;
      dt2 = (da1 / a1 + da2 / a2) * t2;
   } else {
      t1 = 0.0;
      dt1 = 0.0;
      const double a1 = 0.6000000 * exp(0.0570000 * V);
      const double da1 = 0.0570000 * a1;
      const double c = -0.1;
      const double e = exp(c * (V + 32.0000));
      const double a2 = 1.0 / (1.0 + e);
      const double da2 = -a2 * a2 * e * c;
      t2 = a1 * a2;
      dt2 = (da1 / a1 + da2 / a2) * t2;
   }
   *dR = dt1 + dt2;
   return t1 + t2;
}
typedef struct { double x0, x1, c[4]; double (*f)(double, double *); } hermite_rate;
static hermite_rate H_MOD = {-64.20, -23.3, {0, 0, 0, 0}, h_recip};
static hermite_rate J_MOD = {-48.85, -17.6, {0, 0, 0, 0}, j_recip};
/* makeTauRecipParms + hermite (TT06Tau.cc:137-157) */
static void hermite_init(hermite_rate *h)
{
   double d0, d1;
   const double p0 = h->f(h->x0, &d0), p1 = h->f(h->x1, &d1);
   const double dx = h->x1 - h->x0;
   h->c[0] = p0;
   h->c[1] = d0;
   h->c[2] = 3 * (p1 - p0) / (dx * dx) - (2 * d0 + d1) / dx;
   h->c[3] = -2 * (p1 - p0) / (dx * dx * dx) + (d0 + d1) / (dx * dx);
}
static double hermite_rate_eval(const hermite_rate *h, double V)
{
   double d;
   if (V < h->x0 || V > h->x1) return h->f(V, &d);
   const double x = V - h->x0;
   return h->c[0] + x * (h->c[1] + x * (h->c[2] + x * h->c[3]));
}
static double hTauRMod(double V) { return hermite_rate_eval(&H_MOD, V); }
static double jTauRMod(double V) { return hermite_rate_eval(&J_MOD, V); }
static void mod_init(void)
{
   static int done = 0;
   if (done) return;
   hermite_init(&H_MOD);
   hermite_init(&J_MOD);
   done = 1;
}

/* the Hermite cubics' coefficients {x0, x1, c0..c3} for h (which = 0) and
 * j (which = 1) */
void ora_tt06_mod_coefficients(int which, double out[6])
{
   mod_init();
   const hermite_rate *h = which ? &J_MOD : &H_MOD;
   out[0] = h->x0;
   out[1] = h->x1;
   for (int k = 0; k < 4; k++) out[2 + k] = h->c[k];
}

typedef double (*gfun)(double);
static const gfun EXACT_GATE[26] = {
   mMhu, mTauR, hjMhu, hTauR, hjMhu, jTauR, Xr1Mhu, Xr1TauR, Xr2Mhu, Xr2TauR,
   XsMhu, XsTauR, rMhu, rTauR, dMhu, dTauR, fMhu, fTauR, f2Mhu, f2TauR,
   jLMhu, jLTauR, sMhu0, sTauR0, sMhu1, sTauR1
};

/* exact function idx in fit order (0..32); jhTauSmooth selects the smoothed
 * h and j rates for idx 10 and 12, as mkFitParms_ names them
 * (TT06Dev_Reaction.cc:217-219) */
double ora_tt06_exact_func(int idx, double V, int jhTauSmooth)
{
   ora_tt06_init();
   mod_init();
   if (jhTauSmooth && idx == 10) return hTauRMod(V);
   if (jhTauSmooth && idx == 12) return jTauRMod(V);
   switch (idx) {
   case 0: return fv0_x(V);
   case 1: return fv1_x(V);
   case 2: return fv2_x(V);
   case 3: return fv3_x(V);
   case 4: return fv4_x(V);
   case 5: return fv5_x(V);
   case 6: return fv6_x(V);
   default: return EXACT_GATE[idx - 7](V);
   }
}

/* generic updateGate with exact functions (TT06Func.cc:64-161): m by
 * Rush-Larsen with libm exp, the others g += dt*(mhu - g)*tauR; every cell of
 * the group (no v2 loop bounds).  jhTauSmooth = 1 takes the smoothed h and j
 * rates. */
void ora_tt06_gates_exact(double dt, int s_switch, int jhTauSmooth, int n,
                          const double *VM, double *state, int64_t stride)
{
   ora_tt06_init();
   mod_init();
   for (int eq = 0; eq < 12; eq++) {
      int fi = (eq == 11 && s_switch) ? 12 : eq;
      gfun mhuF = EXACT_GATE[2 * fi], tauF = EXACT_GATE[2 * fi + 1];
      if (jhTauSmooth && eq == 1) tauF = hTauRMod;
      if (jhTauSmooth && eq == 2) tauF = jTauRMod;
      double *g = state + (ORA_M + eq) * stride;
      for (int ii = 0; ii < n; ii++) {
         double Vm = VM[ii];
         double mhu = mhuF(Vm), tauR = tauF(Vm);
         if (eq == 0) g[ii] += (mhu - g[ii]) * (1 - exp(-dt * tauR));
         else         g[ii] += dt * (mhu - g[ii]) * tauR;
      }
   }
}
