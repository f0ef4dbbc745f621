// setup_device.hpp -- slab set-up on the device (setup_device.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fgr_stencil.hpp"
#include "fgr_weights.hpp"
#include "hip_util.hpp"

namespace cdh {

// Device buffers of device_slab_weights, kept by the caller across slabs:
// they only grow, so a slab's set-up allocates nothing (no hipMalloc or
// hipFree, which would wait for the device) in the steady state.
struct SlabSetupScratch {
   DevBuf<uint8_t> tissue;      // per box point
   DevBuf<double> sigbox;       // [box point][6]
   DevBuf<double> sigc;         // [cell][6], uploaded by the caller
   DevBuf<uint32_t> rbi;        // remote cells' box indices, uploaded by the caller
   DevBuf<double> rsig;         // remote cells' conductivities [nremote][6]
   DevBuf<float> W;             // mixed form: plain weights [18][S]
   DevBuf<unsigned> wmax;       // mixed form: largest |weight| as float bits
   DevBuf<double> up;           // mixed form: 2^wexp
   void reserve(int64_t nbox, int64_t n, int64_t nremote, int64_t S, bool compact);
};

// FGR weights of a slab from its cells' conductivities (sc.sigc: device
// [n][6] in slab order, already uploaded), asynchronous on `stream`.  bi:
// the slab's S = 64*ntile box indices on the device, offset by b0 (padding
// cells share cell 0's).  Aw non-null: the 18 float weights in the 64-cell
// tile layout; else the compact form into Aref ([ntile][18]) and Aw16 (tile
// layout, 9 words), with 2^-wexp written to the device float *wdelta_scale
// (the slab's SlabDev field), as slab.cpp's compact_weights.  Remote cells
// (nremote; sc.rbi box indices within the slab's box and sc.rsig their
// conductivities) are tissue for the weights.
void device_slab_weights(hipStream_t stream, SlabSetupScratch &sc, const BoxGeom &geom,
                         int64_t nbox, int64_t n, int64_t ntile, const uint32_t *bi, uint32_t b0,
                         int64_t nremote, float *Aw, float *Aref, uint32_t *Aw16,
                         float *wdelta_scale);

// The compact weights' exponent: with the slab's largest |weight| m (a
// float), wexp = 14 - ceil(log2 m), 0 for m = 0.  Written exactly from
// frexp (m = f * 2^e, f in [0.5, 1): ceil(log2 m) = e - 1 for f = 0.5, else
// e), so that the host and the device set-up agree bit for bit.
__host__ __device__ inline int compact_weight_exponent(float m)
{
   if (!(m > 0.0f)) return 0;
   int e = 0;
   const float f = frexpf(m, &e);
   return 14 - (f == 0.5f ? e - 1 : e);
}

// precomputeCoefficients on the device for ncell cells at box points
// bidx[] of a box laid out as G.off describes (device arrays; asynchronous
// on stream): A0[c] and, by layout, Aw[(k-1)*stride + c] (kWeightsPlain),
// Aw[blk_index(k-1, c, 18)] (kWeightsTiled) or Aw[grp_index18(k-1, c)]
// (kWeightsGrouped)
enum WeightLayout : int { kWeightsPlain = 0, kWeightsTiled = 1, kWeightsGrouped = 2 };
void device_fgr_weights(hipStream_t stream, const uint8_t *tissue, const double *sigma,
                        int64_t ncell, const int64_t *bidx, const fgr::WeightGeom &G, double *A0,
                        float *Aw, int64_t stride, int layout = kWeightsPlain);

}  // namespace cdh
