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

void fgr_offsets_strided(int64_t sx, int64_t sy, int64_t sz, int64_t off[19])
{
   for (int k = 0; k < 19; k++)
      off[k] = sx * kNodeDelta[k][0] + sy * kNodeDelta[k][1] + sz * kNodeDelta[k][2];
}

void fgr_offsets(int bx, int by, int bz, int64_t off[19])
{
   (void)bx;
   fgr_offsets_strided((int64_t)by * bz, bz, 1, off);
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
