// fgr_stencil.hpp -- FGR weights of one cell, one code for the host set-up
// (fgr_weights.cpp) and the device set-up kernels (setup_device.hip), so
// that both round alike.  Compile users with -ffp-contract=off.
//
// elec/FGRDiffusionOMP.cc:194-221: for each face whose neighbour is tissue,
// sigma_bar = (sigma_c + sigma_n)/2, sigmaTimesS = (sigma_bar S)/V with S
// the signed face area vector (f1, :234-265), and the face's gradient
// stencil gradPhi (elec/FGRUtils.cc:81-148: +-1 on the normal component;
// +-0.25 on each transverse component, only when its four nodes are tissue,
// :55-121) is accumulated as A[i] += sigmaTimesS[j] gradPhi[j][i] hInv[j]
// into float weights, faces in order.  Within one face every node takes at
// most one nonzero term, so adding only the nonzero terms, in any order
// inside the face, rounds as the full i, j loop does (a zero term leaves a
// float weight unchanged).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "fgr_weights.hpp"

namespace cdh {
namespace fgr {

// NodeLocation order, elec/FGRUtils.hh:19-22
enum : int { ZZZ, MZZ, ZMZ, PZZ, ZPZ, ZZM, ZZP, MMZ, PMZ, PPZ, MPZ, MZM, ZMM, PZM, ZPM, MZP, ZMP,
             PZP, ZPP };

struct WeightGeom {
   int64_t off[19];      // NodeLocation offsets in the box
   int64_t face[6];      // -x, -y, +x, +y, -z, +z
   double h[3], hInv[3], vol;
};

// for a box with NodeLocation offsets off[] (any layout)
inline WeightGeom weight_geom(const BoxGeom &g, const int64_t off[19])
{
   WeightGeom G;
   for (int k = 0; k < 19; k++) G.off[k] = off[k];
   const int fn[6] = {MZZ, ZMZ, PZZ, ZPZ, ZZM, ZZP};
   for (int f = 0; f < 6; f++) G.face[f] = G.off[fn[f]];
   G.h[0] = g.dx;
   G.h[1] = g.dy;
   G.h[2] = g.dz;
   G.hInv[0] = 1.0 / g.dx;
   G.hInv[1] = 1.0 / g.dy;
   G.hInv[2] = 1.0 / g.dz;
   G.vol = G.h[0] * G.h[1] * G.h[2];
   return G;
}

// for the reference's box layout, z fastest
inline WeightGeom weight_geom(const BoxGeom &g)
{
   int64_t off[19];
   fgr_offsets(g.bx, g.by, g.bz, off);
   return weight_geom(g, off);
}

// one term sigmaTimesS[j] * gradPhi[j][i] * hInv[j] into a float weight
__host__ __device__ __forceinline__ void term(float &w, double sS, double g, double hInv)
{
   w += sS * g * hInv;
}

// a transverse component: +0.25 at n0 and n1, -0.25 at n3 and n4
__host__ __device__ __forceinline__ void transverse(float *w, const int *t, double sS, double hInv,
                                                    int n3, int n1, int n4, int n0)
{
   if (t[n4] && t[n3] && t[n0] && t[n1]) {
      term(w[n0], sS, 0.25, hInv);
      term(w[n1], sS, 0.25, hInv);
      term(w[n3], sS, -0.25, hInv);
      term(w[n4], sS, -0.25, hInv);
   }
}

template <int F>
__host__ __device__ __forceinline__ void face(const uint8_t *tissue, const double *sigma,
                                              int64_t ib, const int *t, const WeightGeom &G,
                                              float *w)
{
   const int64_t nb = ib + G.face[F];
   if (!tissue[nb]) return;
   const double *sa = sigma + 6 * ib, *sb = sigma + 6 * nb;
   double s[6];
   for (int q = 0; q < 6; q++) s[q] = (sa[q] + sb[q]) / 2.0;
   constexpr int ax = (F == 0 || F == 2) ? 0 : (F == 1 || F == 3) ? 1 : 2;
   constexpr bool neg = F < 2 || F == 4;
   const double area = ax == 0 ? G.h[1] * G.h[2] : ax == 1 ? G.h[0] * G.h[2] : G.h[0] * G.h[1];
   double S[3] = {0.0, 0.0, 0.0};
   S[ax] = neg ? -area : area;
   double sS[3];
   sS[0] = s[0] * S[0] + s[1] * S[1] + s[2] * S[2];
   sS[1] = s[1] * S[0] + s[3] * S[1] + s[4] * S[2];
   sS[2] = s[2] * S[0] + s[4] * S[1] + s[5] * S[2];
   sS[0] /= G.vol;
   sS[1] /= G.vol;
   sS[2] /= G.vol;
   const double *hi = G.hInv;
   if constexpr (F == 0) {            // -x
      term(w[ZZZ], sS[0], 1.0, hi[0]);
      term(w[MZZ], sS[0], -1.0, hi[0]);
      transverse(w, t, sS[1], hi[1], MMZ, MPZ, ZMZ, ZPZ);
      transverse(w, t, sS[2], hi[2], ZZM, ZZP, MZM, MZP);
   } else if constexpr (F == 1) {     // -y
      transverse(w, t, sS[0], hi[0], MZZ, PZZ, MMZ, PMZ);
      term(w[ZZZ], sS[1], 1.0, hi[1]);
      term(w[ZMZ], sS[1], -1.0, hi[1]);
      transverse(w, t, sS[2], hi[2], ZMM, ZMP, ZZM, ZZP);
   } else if constexpr (F == 2) {     // +x
      term(w[PZZ], sS[0], 1.0, hi[0]);
      term(w[ZZZ], sS[0], -1.0, hi[0]);
      transverse(w, t, sS[1], hi[1], ZMZ, ZPZ, PMZ, PPZ);
      transverse(w, t, sS[2], hi[2], PZM, PZP, ZZM, ZZP);
   } else if constexpr (F == 3) {     // +y
      transverse(w, t, sS[0], hi[0], MPZ, PPZ, MZZ, PZZ);
      term(w[ZPZ], sS[1], 1.0, hi[1]);
      term(w[ZZZ], sS[1], -1.0, hi[1]);
      transverse(w, t, sS[2], hi[2], ZZM, ZZP, ZPM, ZPP);
   } else if constexpr (F == 4) {     // -z
      transverse(w, t, sS[0], hi[0], MZM, PZM, MZZ, PZZ);
      transverse(w, t, sS[1], hi[1], ZMZ, ZPZ, ZMM, ZPM);
      term(w[ZZZ], sS[2], 1.0, hi[2]);
      term(w[ZZM], sS[2], -1.0, hi[2]);
   } else {                           // +z
      transverse(w, t, sS[0], hi[0], MZZ, PZZ, MZP, PZP);
      transverse(w, t, sS[1], hi[1], ZMP, ZPP, ZMZ, ZPZ);
      term(w[ZZP], sS[2], 1.0, hi[2]);
      term(w[ZZZ], sS[2], -1.0, hi[2]);
   }
}

// The 19 float weights of the tissue cell at box point ib; tissue[] and
// sigma[6*] cover the box.
__host__ __device__ __forceinline__ void cell_weights(const uint8_t *tissue, const double *sigma,
                                                      int64_t ib, const WeightGeom &G, float w[19])
{
   int t[19];
   for (int k = 0; k < 19; k++) t[k] = tissue[ib + G.off[k]];
   for (int k = 0; k < 19; k++) w[k] = 0.0f;
   face<0>(tissue, sigma, ib, t, G, w);
   face<1>(tissue, sigma, ib, t, G, w);
   face<2>(tissue, sigma, ib, t, G, w);
   face<3>(tissue, sigma, ib, t, G, w);
   face<4>(tissue, sigma, ib, t, G, w);
   face<5>(tissue, sigma, ib, t, G, w);
}

// A0 = -sum_{i=1..18} A[i], in double, in i order
__host__ __device__ __forceinline__ double centre_weight(const float w[19])
{
   double a0 = 0.0;
   for (int i = 1; i < 19; i++) a0 -= w[i];
   return a0;
}

// Round to the nearest FP16 value (ties to even), as a double; used for the
// compact weights of the mixed form on both sides so they round alike.
// Finite results only: |v| stays below 65520 in that use.
__host__ __device__ __forceinline__ double half_rne(double v)
{
   if (v == 0.0) return v;
   const double a = std::fabs(v);
   int e = std::ilogb(a);
   if (e < -14) e = -14;                         // FP16 subnormals: quantum 2^-24
   const double q = std::ldexp(1.0, e - 10);
   return std::copysign(std::rint(a / q) * q, v);
}

// the FP16 bits of a value half_rne returned (exactly representable)
__host__ __device__ __forceinline__ uint16_t half_bits(double exact)
{
   const _Float16 h = (_Float16)(float)exact;
   uint16_t b;
   __builtin_memcpy(&b, &h, 2);
   return b;
}

}  // namespace fgr
}  // namespace cdh
