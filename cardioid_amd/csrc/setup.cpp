// setup.cpp -- host-side set-up shared by the drop-in objects and the engine.
#include "setup.hpp"

#include <cmath>
#include <string>

#include "hip_util.hpp"

namespace cdh {

void check_configuration_b(double tolerance, int mod, int fastGate, int fastNonGate)
{
   if (!(std::fabs(tolerance / 1e-4 - 1.0) < 1e-12) || mod != 1)
      throw Error(kErrUnsupported,
                  "TT06Dev: only tolerance = 1e-4 with mod = 1 (rational fits) is implemented");
   if (fastGate < 1)
      throw Error(kErrUnsupported,
                  "TT06Dev: fastGate = 0 selects the generic gate path, which applies dt twice "
                  "with rational fits; use fastGate >= 1 (the v2 gate kernels)");
   if (fastNonGate != 0)
      throw Error(kErrUnsupported,
                  "TT06Dev: fastNonGate > 0 (approximate SIMD non-gates) is not implemented");
}

std::vector<double> scaled_fit_coefficients(const int *l, const int *m, const double *coef,
                                            double dt)
{
   CDH_REQUIRE(l && m && coef, "fit set: null argument");
   CDH_REQUIRE(dt > 0, "fit set: dt must be positive");
   std::vector<double> out(kFitCoefTotal, 0.0);
   int pos = 0;
   for (int i = 0; i < 33; i++) {
      if (l[i] != kFitOrder[i].l || m[i] != kFitOrder[i].m)
         throw Error(kErrUnsupported, "fit " + std::to_string(i) + " has orders (l=" +
                                          std::to_string(l[i]) + ", m=" + std::to_string(m[i]) +
                                          "); the kernels are built for (l=" +
                                          std::to_string(kFitOrder[i].l) + ", m=" +
                                          std::to_string(kFitOrder[i].m) + ")");
      const bool tauR = i >= 7 && ((i - 7) % 2 == 1);
      for (int j = 0; j < l[i] + m[i]; j++) {
         double c = coef[pos + j];
         if (tauR && j < m[i]) c *= dt;
         out[fit_offset(i) + j] = c;
      }
      pos += l[i] + m[i];
   }
   return out;
}

ParamSet make_param_set(const CellType &ct, double s_kr, double s_cal, double s_na)
{
   TT06Const K;
   tt06_constants(&K);
   ParamSet p;
   p.P_NaK = ct.P_NaK;
   p.g_Ks = ct.g_Ks;
   p.c_Kr = s_kr * ct.g_Kr;
   p.g_to = ct.g_to;
   p.g_NaL = ct.g_NaL;
   p.c_CaL = s_cal;
   p.c_Na = s_na * K.c20;
   p.K0 = ct.state0[S_DVK_I];
   p.Vm0 = ct.Vm0;
   p.s_switch = ct.s_switch;
   p.pad_ = 0;
   return p;
}

std::vector<double> rest_state_soa(const CellType &ct, int64_t n, int64_t stride, double c9)
{
   std::vector<double> st((size_t)N_STATE * stride, 0.0);
   for (int s = 0; s < N_STATE; s++)
      for (int64_t i = 0; i < stride; i++) st[s * stride + i] = ct.state0[s];
   for (int64_t i = 0; i < stride; i++) st[S_DVK_I * stride + i] = ct.state0[S_DVK_I] / c9 + ct.Vm0;
   (void)n;
   return st;
}

BoxGeom bounding_box(int nx, int ny, int nz, const int64_t *gid, int64_t nLocal, double dx,
                     double dy, double dz)
{
   (void)nz;
   BoxGeom g;
   g.dx = dx;
   g.dy = dy;
   g.dz = dz;
   if (nLocal <= 0) {
      g.bx = g.by = g.bz = 0;
      g.x0 = g.y0 = g.z0 = 0;
      return g;
   }
   int64_t xmin = INT64_MAX, ymin = INT64_MAX, zmin = INT64_MAX, xmax = -1, ymax = -1, zmax = -1;
   for (int64_t c = 0; c < nLocal; c++) {
      const int64_t x = gid[c] % nx, y = (gid[c] / nx) % ny, z = gid[c] / ((int64_t)nx * ny);
      xmin = std::min(xmin, x); xmax = std::max(xmax, x);
      ymin = std::min(ymin, y); ymax = std::max(ymax, y);
      zmin = std::min(zmin, z); zmax = std::max(zmax, z);
   }
   g.bx = (int)(xmax - xmin + 3);
   g.by = (int)(ymax - ymin + 3);
   g.bz = (int)(zmax - zmin + 3);
   g.x0 = (int)xmin - 1;
   g.y0 = (int)ymin - 1;
   g.z0 = (int)zmin - 1;
   return g;
}

int64_t box_index(const BoxGeom &g, int nx, int ny, int64_t gid)
{
   const int64_t x = gid % nx - g.x0, y = (gid / nx) % ny - g.y0, z = gid / ((int64_t)nx * ny) - g.z0;
   if (x < 0 || y < 0 || z < 0 || x >= g.bx || y >= g.by || z >= g.bz) return -1;
   return z + (int64_t)g.bz * (y + (int64_t)g.by * x);
}

}  // namespace cdh
