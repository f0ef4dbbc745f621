/*
 * fgr_oracle.c -- CPU restatement of the FGR 19-point diffusion (TEST INFRASTRUCTURE ONLY).
 *
 * The fibre conductivity follows elec/FibreConductivity.cc:27-62
 * (calcConductivityMatrixIBT): with s/c the sines and cosines of phi and
 * theta, a11 = sL c_p^2 s_t^2 + sT (c_p^2 c_t^2 + s_p^2), a22 = sL s_p^2 s_t^2
 * + sT (c_p^2 + s_p^2 c_t^2), a33 = sL c_t^2 + sT s_t^2, a12 = (sL - sT) c_p
 * s_t^2 s_p, a13 = (sL - sT) c_t c_p s_t, a23 = (sL - sT) s_p c_t s_t.
 *
 * Weights follow elec/FGRDiffusionOMP.cc:166-265 and elec/FGRUtils.cc:10-148:
 * per tissue cell and per tissue face neighbour, sigma_bar = (sigma_c +
 * sigma_n)/2, sigmaTimesS = (sigma_bar * S)/V, and
 *     A[i] += sigmaTimesS[j] * gradPhi[j][i] * hInv[j]      (i = 0..18, j = 0..2)
 * accumulated into float weights (one rounding per +=), then
 *     A0 = -sum_{i=1..18} A[i]   in double.
 * The stencil (FGRDiffusionOMP.cc:117-128) is
 *     dVm = scale * (A0*phi0 + sum_{k=1..18} A[k]*phi_k)   in double, in k order.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

/* NodeLocation order, elec/FGRUtils.hh:19-22 */
enum { ZZZ, MZZ, ZMZ, PZZ, ZPZ, ZZM, ZZP, MMZ, PMZ, PPZ, MPZ, MZM, ZMM, PZM, ZPM, MZP, ZMP, PZP, ZPP };

static const int NODE_D[19][3] = {
   {0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
   {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, 0, -1}, {0, -1, -1},
   {1, 0, -1}, {0, 1, -1}, {-1, 0, 1}, {0, -1, 1}, {1, 0, 1}, {0, 1, 1}};

void ora_fgr_offsets(int bx, int by, int bz, int64_t off[19])
{
   (void)bx;
   for (int k = 0; k < 19; k++)
      off[k] = (int64_t)NODE_D[k][2] + (int64_t)bz * ((int64_t)NODE_D[k][1] + (int64_t)by * NODE_D[k][0]);
}

/* setGradientWeights: only case 15 (all four tangential nodes tissue) is
 * non-zero, with w0 = w1 = 0.25 and w3 = w4 = -0.25 (FGRUtils.cc:55-121). */
static void grad_w(double *grad, const int *t, int n3, int n2, int n1, int n4, int n5, int n0)
{
   (void)n2; (void)n5;
   int c = t[n4] + 2 * t[n3] + 4 * t[n0] + 8 * t[n1];
   if (c == 15) {
      grad[n0] = grad[n1] = 0.25;
      grad[n3] = grad[n4] = -0.25;
   }
}

void ora_fgr_f2(unsigned iFace, const int tissue[19], double g[3][19])
{
   switch (iFace) {
   case 0:
      g[0][ZZZ] = 1.0; g[0][MZZ] = -1.0;
      grad_w(g[1], tissue, MMZ, MZZ, MPZ, ZMZ, ZZZ, ZPZ);
      grad_w(g[2], tissue, ZZM, ZZZ, ZZP, MZM, MZZ, MZP);
      break;
   case 1:
      grad_w(g[0], tissue, MZZ, ZZZ, PZZ, MMZ, ZMZ, PMZ);
      g[1][ZZZ] = 1.0; g[1][ZMZ] = -1.0;
      grad_w(g[2], tissue, ZMM, ZMZ, ZMP, ZZM, ZZZ, ZZP);
      break;
   case 2:
      g[0][PZZ] = 1.0; g[0][ZZZ] = -1.0;
      grad_w(g[1], tissue, ZMZ, ZZZ, ZPZ, PMZ, PZZ, PPZ);
      grad_w(g[2], tissue, PZM, PZZ, PZP, ZZM, ZZZ, ZZP);
      break;
   case 3:
      grad_w(g[0], tissue, MPZ, ZPZ, PPZ, MZZ, ZZZ, PZZ);
      g[1][ZPZ] = 1.0; g[1][ZZZ] = -1.0;
      grad_w(g[2], tissue, ZZM, ZZZ, ZZP, ZPM, ZPZ, ZPP);
      break;
   case 4:
      grad_w(g[0], tissue, MZM, ZZM, PZM, MZZ, ZZZ, PZZ);
      grad_w(g[1], tissue, ZMZ, ZZZ, ZPZ, ZMM, ZZM, ZPM);
      g[2][ZZZ] = 1.0; g[2][ZZM] = -1.0;
      break;
   case 5:
      grad_w(g[0], tissue, MZZ, ZZZ, PZZ, MZP, ZZP, PZP);
      grad_w(g[1], tissue, ZMP, ZZP, ZPP, ZMZ, ZZZ, ZPZ);
      g[2][ZZP] = 1.0; g[2][ZZZ] = -1.0;
      break;
   }
}

void ora_fgr_weights(int bx, int by, int bz, const uint8_t *tissue,
                     const double *sigma, double dx, double dy, double dz,
                     int64_t ncell, const int64_t *blockIndex, float *A, double *A0)
{
   int64_t off[19];
   ora_fgr_offsets(bx, by, bz, off);
   const int64_t faceOff[6] = {off[MZZ], off[ZMZ], off[PZZ], off[ZPZ], off[ZZM], off[ZZP]};
   const double h[3] = {dx, dy, dz};
   const double hInv[3] = {1.0 / dx, 1.0 / dy, 1.0 / dz};
   const double vol = h[0] * h[1] * h[2];

#pragma omp parallel for schedule(static)
   for (int64_t c = 0; c < ncell; c++) {
      const int64_t ib = blockIndex[c];
      float w[19];
      for (int k = 0; k < 19; k++) w[k] = 0.0f;
      int t[19];
      for (int k = 0; k < 19; k++) t[k] = tissue[ib + off[k]];
      for (int f = 0; f < 6; f++) {
         const int64_t nb = ib + faceOff[f];
         if (tissue[nb] == 0) continue;
         /* f1: face-averaged tensor times the signed face area */
         const double *sa = sigma + 6 * ib, *sb = sigma + 6 * nb;
         double s11 = (sa[0] + sb[0]) / 2.0, s12 = (sa[1] + sb[1]) / 2.0, s13 = (sa[2] + sb[2]) / 2.0;
         double s22 = (sa[3] + sb[3]) / 2.0, s23 = (sa[4] + sb[4]) / 2.0, s33 = (sa[5] + sb[5]) / 2.0;
         double S[3] = {0.0, 0.0, 0.0};
         switch (f) {
         case 0: S[0] = -h[1] * h[2]; break;
         case 1: S[1] = -h[0] * h[2]; break;
         case 2: S[0] = h[1] * h[2]; break;
         case 3: S[1] = h[0] * h[2]; break;
         case 4: S[2] = -h[0] * h[1]; break;
         case 5: S[2] = h[0] * h[1]; break;
         }
         double sS[3];
         sS[0] = s11 * S[0] + s12 * S[1] + s13 * S[2];
         sS[1] = s12 * S[0] + s22 * S[1] + s23 * S[2];
         sS[2] = s13 * S[0] + s23 * S[1] + s33 * S[2];
         sS[0] /= vol; sS[1] /= vol; sS[2] /= vol;
         double gp[3][19];
         memset(gp, 0, sizeof(gp));
         ora_fgr_f2((unsigned)f, t, gp);
         for (int i = 0; i < 19; i++)
            for (int j = 0; j < 3; j++)
               w[i] += sS[j] * gp[j][i] * hInv[j];
      }
      double a0 = 0.0;
      for (int i = 1; i < 19; i++) a0 -= w[i];
      for (int k = 0; k < 19; k++) A[19 * c + k] = w[k];
      A0[c] = a0;
   }
}

void ora_fibre_conductivity(double theta, double phi, double sigmaL, double sigmaT,
                            double out[6])
{
   const double sp = sin(phi), cp = cos(phi), st = sin(theta), ct = cos(theta);
   const double sp2 = sp * sp, cp2 = cp * cp, st2 = st * st, ct2 = ct * ct;
   out[0] = (sigmaL * cp2 * st2) + (sigmaT * ((cp2 * ct2) + sp2));
   out[1] = (sigmaL - sigmaT) * cp * st2 * sp;
   out[2] = (sigmaL - sigmaT) * ct * cp * st;
   out[3] = (sigmaL * sp2 * st2) + (sigmaT * (cp2 + (sp2 * ct2)));
   out[4] = (sigmaL - sigmaT) * sp * ct * st;
   out[5] = (sigmaL * ct2) + (sigmaT * st2);
}

void ora_fgr_calc(int64_t ncell, const int64_t *blockIndex, const float *A,
                  const double *A0, const int64_t off[19], const double *VmBlock,
                  double scale, double *dVm)
{
#pragma omp parallel for schedule(static)
   for (int64_t c = 0; c < ncell; c++) {
      const double *phi = VmBlock + blockIndex[c];
      const float *a = A + 19 * c;
      double tmp = A0[c] * phi[off[0]];
      for (int k = 1; k < 19; k++) tmp += a[k] * phi[off[k]];
      dVm[c] = tmp * scale;
   }
}
