// tt06_host.cpp -- host-side TT06Dev setup: derived constants and the
// standard cell types.
//
// The derived constants follow elec/TT06NonGates.c:137-208 expression by
// expression (same operand order), so that the doubles handed to the
// kernels are the ones the reference computes.
#include <cmath>
#include <cstring>

#include "tt06_model.hpp"

namespace cdh {

void tt06_constants(TT06Const *k)
{
   // TT06 (2006) physical parameters (the CellML values the reference uses)
   const double R = 8314.472, T = 310, F = 96485.3415, Cm = 0.185, V_c = 0.016404;
   const double P_kna = 0.03, K_o = 5.4, Na_o = 140, Ca_o = 2, g_K1 = 5.405;
   const double g_Na = 14.838, g_bna = 0.00029, g_bca = 0.000592, g_pCa = 0.1238;
   const double K_mNa = 40, K_pCa = 0.0005;
   const double k1p = 0.15, k2p = 0.045, k3 = 0.06, k4 = 0.005, EC = 1.5;
   const double max_sr = 2.5, min_sr = 1, V_rel = 0.102, V_xfer = 0.0038, K_up = 0.00025;
   const double V_leak = 0.00036, Vmax_up = 0.006375, Buf_c = 0.2, K_buf_c = 0.001;
   const double Buf_sr = 10, K_buf_sr = 0.3, Buf_ss = 0.4, K_buf_ss = 0.00025;
   const double V_sr = 0.001094, V_ss = 0.00005468;

   const double c1 = F / (R * T);
   k->c2 = P_kna;
   k->c3 = -1 / c1;
   k->c4 = -k->c3 * std::log(Na_o);
   k->c5 = -k->c3 * std::log(K_o);
   k->c6 = -k->c3 * std::log(K_o + P_kna * Na_o);
   k->c8 = -0.5 * k->c3 * std::log(Ca_o);
   k->c9 = -Cm / (V_c * F);
   k->c7 = g_bca;
   k->c11 = std::sqrt(K_o / 5.4);
   k->c12 = g_K1 * std::sqrt(K_o / 5.4);
   k->c13 = Cm / (2.0 * V_ss * F * k->c9);
   k->c14 = V_sr / V_ss;
   k->c15 = -V_ss / V_c;
   k->c16 = V_sr / V_c;
   k->c17 = k3 / (k1p * k2p) * k->c9;
   k->c18 = k2p * max_sr / k->c9;
   k->c19 = -k2p * (max_sr - min_sr) / k->c9;
   k->c20 = g_Na;
   k->c21 = g_bna;
   k->c22 = 1 / k->c9;
   k->c23 = V_xfer / (k->c15 * k->c9);
   k->c24 = g_pCa;
   k->c25 = 1.0 / K_mNa;
   k->c26 = 1.0 / K_pCa;
   k->c27 = 1.0 / K_up;
   k->c28 = 1.0 / std::sqrt(Buf_c * K_buf_c);
   k->c29 = K_buf_c * k->c28;
   k->c30 = 1.0 / EC;
   k->c31 = 1.0 / std::sqrt(Buf_sr * K_buf_sr);
   k->c32 = K_buf_sr * k->c31;
   k->c33 = 1.0 / std::sqrt(Buf_ss * K_buf_ss);
   k->c34 = K_buf_ss * k->c33;
   k->c36 = k4 / k->c9;
   k->c40 = V_rel / k->c9;
   k->c43 = V_leak / k->c9;
   k->c44 = Vmax_up / k->c9;
   k->exp_coef[0] = 1.0;
   for (int i = 1; i < 17; i++) k->exp_coef[i] = k->exp_coef[i - 1] / i;
}


namespace {

// 1/tau of the h and j gates with its derivative, the two branches of
// hTauRecip / jTauRecip (elec/TT06Tau.cc:14-129): only the value and first
// derivative at the bridge's two ends enter the cubic, in the reference's
// operation order so that the coefficients are its doubles.
double h_rate(double V, double *d)
{
   double t1, t2, d1, d2;
   if (V < -40.0) {
      t1 = 0.0570000 * std::exp(-(V + 80.0) / 6.80000);
      d1 = -t1 / 6.8000;
      t2 = 2.70000 * std::exp(0.0790000 * V) + 310000. * std::exp(0.348500 * V);
      d2 = 0.079 * 2.70000 * std::exp(0.0790000 * V) + 0.3485 * 310000. * std::exp(0.348500 * V);
   } else {
      const double c = (0.770000 / 0.130000);
      const double e = std::exp((V + 10.6600) / -11.1000);
      const double s = 1.0 / (1.0 + e);
      t1 = 0.0;
      d1 = 0.0;
      t2 = c * s;
      d2 = c * s * s * e / 11.1;
   }
   *d = d1 + d2;
   return t1 + t2;
}

double j_rate(double V, double *d)
{
   double t1, t2, d1, d2;
   if (V < -40.0) {
      const double e1 = std::exp(0.244400 * V), e2 = std::exp(-0.0439100 * V);
      const double a1 = (-25428.0 * e1 - 6.94800e-06 * e2);
      const double da1 = (-0.24400 * 25428.0 * e1 + 0.0439100 * 6.94800e-06 * e2);
      const double a2 = (V + 37.7800), da2 = 1.0;
      const double e3 = std::exp(0.311 * (V + 79.2300));
      const double a3 = 1 / (1.0 + e3);
      const double da3 = -a3 * a3 * e3 * 0.311;
      t1 = a1 * a2 * a3;
      d1 = (da1 / a1 + da2 / a2 + da3 / a3) * t1;
      const double b1 = 0.0242400 * std::exp(-0.0105200 * V);
      const double db1 = -0.0105200 * b1;
      const double e4 = std::exp(-0.137800 * (V + 40.1400));
      const double b2 = 1.0 / (1.0 + e4);
      const double db2 = -(-0.137800) * b2 * b2 * e4;
      t2 = b1 * b2;
      d2 = (db1 / b1 + db2 / b2) * t2;
   } else {
      t1 = 0.0;
      d1 = 0.0;
      const double b1 = 0.6000000 * std::exp(0.0570000 * V);
      const double db1 = 0.0570000 * b1;
      const double e = std::exp(-0.1 * (V + 32.0000));
      const double b2 = 1.0 / (1.0 + e);
      const double db2 = -b2 * b2 * e * -0.1;
      t2 = b1 * b2;
      d2 = (db1 / b1 + db2 / b2) * t2;
   }
   *d = d1 + d2;
   return t1 + t2;
}

// makeTauRecipParms + hermite (elec/TT06Tau.cc:137-157)
void bridge(double x0, double x1, double (*rate)(double, double *), double out[6])
{
   double d0, d1;
   const double p0 = rate(x0, &d0), p1 = rate(x1, &d1);
   const double dx = x1 - x0;
   out[0] = x0;
   out[1] = x1;
   out[2] = p0;
   out[3] = d0;
   out[4] = 3 * (p1 - p0) / (dx * dx) - (2 * d0 + d1) / dx;
   out[5] = -2 * (p1 - p0) / (dx * dx * dx) + (d0 + d1) / (dx * dx);
}

struct Row {
   const char *name;
   int s_switch;
   double P_NaK, g_Ks, g_Kr, g_to, g_NaL, Vm;
   double s[N_STATE];
};
// elec/TT06Func.cc:421-612 (rest states in slot order Ca_i, K_i, Na_i, Ca_ss,
// Ca_SR, R', fCass, m, h, j, Xr1, Xr2, Xs, r, d, f, f2, jL, s)
const Row kRows[] = {
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
}  // namespace

// The exact functions' constants (elec/TT06NonGates.c:236-248, same
// operand order) and, for mod = 1, the h and j rate bridges
// (TT06Func::initCnst, elec/TT06Func.cc:56-61).
void tt06_exact_constants(TT06ExactConst *e, int mod)
{
   const double R = 8314.472, T = 310, F = 96485.3415;
   const double K_o = 5.4, Na_o = 140, Ca_o = 2, g_CaL = 0.0000398, g_pK = 0.0146;
   const double K_mk = 1, K_NaCa = 1000, K_sat = 0.1, alpha = 2.5, gamma_ = 0.35;
   const double Km_Ca = 1.38, Km_Nai = 87.5;
   const double c1 = F / (R * T);
   e->f1 = c1;
   e->f2 = -2.0 * K_o / (K_o + K_mk);
   e->f3 = ((Km_Nai * Km_Nai * Km_Nai + Na_o * Na_o * Na_o) * (Km_Ca + Ca_o)) / (K_NaCa * Ca_o);
   e->f4 = e->f3 * K_sat;
   e->f5 = gamma_ * e->f1;
   e->f6 = (Na_o * Na_o * Na_o * alpha / Ca_o);
   e->f7a = 0.5 * g_CaL * F;
   e->f9a = 4.0 * Ca_o * e->f7a;
   e->f10 = g_pK;
   bridge(-64.20, -23.3, h_rate, e->hmod);
   bridge(-48.85, -17.6, j_rate, e->jmod);
   e->mod = mod ? 1 : 0;
   e->pad_ = 0;
}

bool tt06_standard_celltype(const char *name, CellType *out)
{
   for (const Row &r : kRows) {
      if (std::strcmp(r.name, name) != 0) continue;
      out->s_switch = r.s_switch;
      out->P_NaK = r.P_NaK;
      out->g_Ks = r.g_Ks;
      out->g_Kr = r.g_Kr;
      out->g_to = r.g_to;
      out->g_NaL = r.g_NaL;
      out->Vm0 = r.Vm;
      std::memcpy(out->state0, r.s, sizeof(r.s));
      return true;
   }
   return false;
}

}  // namespace cdh
