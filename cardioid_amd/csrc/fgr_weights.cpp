// fgr_weights.cpp -- FGR stencil weights on the host.
//
// Per tissue cell, the weights of fgr_stencil.hpp (elec/FGRDiffusionOMP.cc:
// 166-265, elec/FGRUtils.cc:10-148) accumulated into float weights; the
// centre weight is minus the sum of the 18 others, in double.  Same rounding
// as the reference's CMake build (float weights, elec/FGRUtils.hh:7-15).
#include "fgr_weights.hpp"
#include "fgr_stencil.hpp"

namespace cdh {

namespace {
const int kNodeDelta[19][3] = {
   {0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
   {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, 0, -1}, {0, -1, -1},
   {1, 0, -1}, {0, 1, -1}, {-1, 0, 1}, {0, -1, 1}, {1, 0, 1}, {0, 1, 1}};
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
   const fgr::WeightGeom G = fgr::weight_geom(g);
#pragma omp parallel for schedule(static)
   for (int64_t c = 0; c < ncell; c++) {
      float w[19];
      fgr::cell_weights(tissue, sigma, blockIndex[c], G, w);
      A0[c] = fgr::centre_weight(w);
      for (int k = 1; k < 19; k++) Aw[(k - 1) * stride + c] = w[k];
   }
}

}  // namespace cdh
