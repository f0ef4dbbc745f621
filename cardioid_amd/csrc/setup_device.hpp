// setup_device.hpp -- slab set-up on the device (setup_device.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fgr_stencil.hpp"
#include "fgr_weights.hpp"

namespace cdh {

// FGR weights of a slab from its cells' conductivities (sig_cells: host
// [n][6] in slab order), on `stream`, synchronously.  bi: the slab's S =
// 64*ntile box indices on the device, offset by b0 (padding cells share
// cell 0's).  Aw non-null: the 18 float weights in the 64-cell tile layout;
// else the compact form into Aref ([ntile][18]) and Aw16 (tile layout, 9
// words) with *wdelta_scale = 2^-wexp, as slab.cpp's compact_weights.
// Remote cells (nremote, host box indices within the slab's box and their
// conductivities [nremote][6]) are tissue for the weights.
void device_slab_weights(hipStream_t stream, const BoxGeom &geom, int64_t nbox, int64_t n,
                         int64_t ntile, const double *sig_cells, const uint32_t *bi, uint32_t b0,
                         int64_t nremote, const uint32_t *remote_bi, const double *remote_sig,
                         float *Aw, float *Aref, uint32_t *Aw16, float *wdelta_scale);

// precomputeCoefficients on the device for ncell cells at box points
// bidx[] of a box laid out as G.off describes (device arrays; asynchronous
// on stream): A0[c] and Aw[(k-1)*stride + c], or with `blocked` the 64-cell
// tile layout Aw[blk_index(k-1, c, 18)]
void device_fgr_weights(hipStream_t stream, const uint8_t *tissue, const double *sigma,
                        int64_t ncell, const int64_t *bidx, const fgr::WeightGeom &G, double *A0,
                        float *Aw, int64_t stride, bool blocked = false);

}  // namespace cdh
