// fgr_weights.cpp -- FGR stencil weights on the host.
//
// Per tissue cell and per tissue face neighbour (elec/FGRDiffusionOMP.cc:
// 194-221): the face-averaged conductivity times the signed face area over
// the cell volume (f1, :234-265), times the face's gradient stencil
// (elec/FGRUtils.cc:81-148, transverse terms only when all four transverse
// nodes are tissue, :55-121), times 1/h, accumulated into float weights
// in the order face -> node -> component; the centre weight is minus the
// sum of the 18 others, in double.  Same rounding as the reference's
// CMake build (float weights, elec/FGRUtils.hh:7-15).
#include <cstring>
#include <vector>

#include "fgr_weights.hpp"

namespace cdh {

namespace {
enum { ZZZ, MZZ, ZMZ, PZZ, ZPZ, ZZM, ZZP, MMZ, PMZ, PPZ, MPZ, MZM, ZMM, PZM, ZPM, MZP, ZMP, PZP, ZPP };

const int kNodeDelta[19][3] = {
   {0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
   {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, 0, -1}, {0, -1, -1},
   {1, 0, -1}, {0, 1, -1}, {-1, 0, 1}, {0, -1, 1}, {1, 0, 1}, {0, 1, 1}};

// transverse gradient weights on one face: nonzero only if the four
// transverse nodes are tissue
inline void transverse(double *g, const int *t, int n3, int n1, int n4, int n0)
{
   if (t[n4] && t[n3] && t[n0] && t[n1]) {
      g[n0] = g[n1] = 0.25;
      g[n3] = g[n4] = -0.25;
   }
}

// gradient stencils of face f (components x, y, z over the 19 nodes)
void face_gradient(int f, const int *t, double g[3][19])
{
   std::memset(g, 0, sizeof(double) * 57);
   switch (f) {
   case 0:  // -x
      g[0][ZZZ] = 1.0; g[0][MZZ] = -1.0;
      transverse(g[1], t, MMZ, MPZ, ZMZ, ZPZ);
      transverse(g[2], t, ZZM, ZZP, MZM, MZP);
      break;
   case 1:  // -y
      transverse(g[0], t, MZZ, PZZ, MMZ, PMZ);
      g[1][ZZZ] = 1.0; g[1][ZMZ] = -1.0;
      transverse(g[2], t, ZMM, ZMP, ZZM, ZZP);
      break;
   case 2:  // +x
      g[0][PZZ] = 1.0; g[0][ZZZ] = -1.0;
      transverse(g[1], t, ZMZ, ZPZ, PMZ, PPZ);
      transverse(g[2], t, PZM, PZP, ZZM, ZZP);
      break;
   case 3:  // +y
      transverse(g[0], t, MPZ, PPZ, MZZ, PZZ);
      g[1][ZPZ] = 1.0; g[1][ZZZ] = -1.0;
      transverse(g[2], t, ZZM, ZZP, ZPM, ZPP);
      break;
   case 4:  // -z
      transverse(g[0], t, MZM, PZM, MZZ, PZZ);
      transverse(g[1], t, ZMZ, ZPZ, ZMM, ZPM);
      g[2][ZZZ] = 1.0; g[2][ZZM] = -1.0;
      break;
   default:  // +z
      transverse(g[0], t, MZZ, PZZ, MZP, PZP);
      transverse(g[1], t, ZMP, ZPP, ZMZ, ZPZ);
      g[2][ZZP] = 1.0; g[2][ZZZ] = -1.0;
      break;
   }
}
}  // namespace

void fgr_offsets(int bx, int by, int bz, int64_t off[19])
{
   (void)bx;
   for (int k = 0; k < 19; k++)
      off[k] = (int64_t)kNodeDelta[k][2] +
               (int64_t)bz * ((int64_t)kNodeDelta[k][1] + (int64_t)by * kNodeDelta[k][0]);
}

void fgr_weights(const BoxGeom &g, const uint8_t *tissue, const double *sigma, int64_t ncell,
                 const int64_t *blockIndex, double *A0, float *Aw, int64_t stride)
{
   int64_t off[19];
   fgr_offsets(g.bx, g.by, g.bz, off);
   const int64_t face[6] = {off[MZZ], off[ZMZ], off[PZZ], off[ZPZ], off[ZZM], off[ZZP]};
   const double h[3] = {g.dx, g.dy, g.dz};
   const double hInv[3] = {1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz};
   const double vol = h[0] * h[1] * h[2];

#pragma omp parallel for schedule(static)
   for (int64_t c = 0; c < ncell; c++) {
      const int64_t ib = blockIndex[c];
      int t[19];
      for (int k = 0; k < 19; k++) t[k] = tissue[ib + off[k]];
      float w[19];
      for (int k = 0; k < 19; k++) w[k] = 0.0f;
      for (int f = 0; f < 6; f++) {
         const int64_t nb = ib + face[f];
         if (!tissue[nb]) continue;
         const double *sa = sigma + 6 * ib, *sb = sigma + 6 * nb;
         double s[6];
         for (int q = 0; q < 6; q++) s[q] = (sa[q] + sb[q]) / 2.0;
         // sigma_bar * S with S the signed face area vector
         double S[3] = {0.0, 0.0, 0.0};
         const int ax = (f == 0 || f == 2) ? 0 : (f == 1 || f == 3) ? 1 : 2;
         const double sgn = (f < 2 || f == 4) ? -1.0 : 1.0;
         const double area = ax == 0 ? h[1] * h[2] : ax == 1 ? h[0] * h[2] : h[0] * h[1];
         S[ax] = sgn < 0 ? -area : area;
         double sS[3];
         sS[0] = s[0] * S[0] + s[1] * S[1] + s[2] * S[2];
         sS[1] = s[1] * S[0] + s[3] * S[1] + s[4] * S[2];
         sS[2] = s[2] * S[0] + s[4] * S[1] + s[5] * S[2];
         sS[0] /= vol;
         sS[1] /= vol;
         sS[2] /= vol;
         double gp[3][19];
         face_gradient(f, t, gp);
         for (int i = 0; i < 19; i++)
            for (int j = 0; j < 3; j++) w[i] += sS[j] * gp[j][i] * hInv[j];
      }
      double a0 = 0.0;
      for (int i = 1; i < 19; i++) a0 -= w[i];
      A0[c] = a0;
      for (int k = 1; k < 19; k++) Aw[(k - 1) * stride + c] = w[k];
   }
}

}  // namespace cdh
