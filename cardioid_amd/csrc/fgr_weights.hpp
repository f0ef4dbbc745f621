// fgr_weights.hpp -- FGR stencil weights (host).
#pragma once

#include <cstdint>

namespace cdh {

// Padded box, z fastest: index = z + bz*(y + by*x) (elec/Array3d.hh:173-177).
struct BoxGeom {
   int bx, by, bz;       // box dimensions
   int x0, y0, z0;       // global tuple of box point (0,0,0)
   double dx, dy, dz;
};

// NodeLocation offsets (elec/FGRUtils.hh:19-22) in the box.
void fgr_offsets(int bx, int by, int bz, int64_t off[19]);
// The same for a box whose x, y and z neighbours lie sx, sy and sz apart.
void fgr_offsets_strided(int64_t sx, int64_t sy, int64_t sz, int64_t off[19]);

// Weights for ncell cells at box indices blockIndex[]: A0[c] (double) and
// Aw[(k-1)*stride + c] (float, k = 1..18).  tissue[] and sigma[6*] are over
// the whole box.
void fgr_weights(const BoxGeom &g, const uint8_t *tissue, const double *sigma, int64_t ncell,
                 const int64_t *blockIndex, double *A0, float *Aw, int64_t stride);

}  // namespace cdh
