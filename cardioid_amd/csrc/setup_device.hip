// setup_device.hip -- set-up on the device: fibre conductivity from angles
// and FGR weights from the conductivity and tissue boxes, with the host's
// rounding (fgr_stencil.hpp; built with -ffp-contract=off, and FP64 division
// on the device is correctly rounded as on the host).
//
//  * k_fibre_sigma: calcConductivityMatrixIBT (elec/FibreConductivity.cc:
//    27-62) per cell.  From integer angles (the pio anatomy's intType
//    branch, elec/setConductivity.cc:184-190: angle = i*pi/256) the sines
//    and cosines come from a host table, so the tensors equal the host's
//    bit for bit; from double angles the device's sin/cos are used.
//  * k_fgr_weights: precomputeCoefficients (elec/FGRDiffusionOMP.cc:166-224)
//    per cell, A0 and the 18 float weights in [k][stride] layout.
//  * The engine's slab set-up: sigma scattered into the slab's box, weights
//    in the arena's 64-cell tile layout, and for the mixed form the compact
//    weights (per-tile FP16-exact reference, per-cell FP16 difference) as
//    slab.cpp builds them.
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "fgr_stencil.hpp"
#include "hip_util.hpp"
#include "setup_device.hpp"
#include "tt06_model.hpp"

namespace cdh {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ void ibt(double sp, double cp, double st, double ct, double sL,
                                    double sT, double *o)
{
   const double sp2 = sp * sp, cp2 = cp * cp, st2 = st * st, ct2 = ct * ct;
   o[0] = (sL * cp2 * st2) + (sT * ((cp2 * ct2) + sp2));
   o[1] = (sL - sT) * cp * st2 * sp;
   o[2] = (sL - sT) * ct * cp * st;
   o[3] = (sL * sp2 * st2) + (sT * (cp2 + (sp2 * ct2)));
   o[4] = (sL - sT) * sp * ct * st;
   o[5] = (sL * ct2) + (sT * st2);
}

__global__ __launch_bounds__(kBlock) void k_fibre_sigma(int64_t n, const double *theta,
                                                        const double *phi, double sL, double sT,
                                                        double *sigma)
{
   const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
   if (c >= n) return;
   ibt(sin(phi[c]), cos(phi[c]), sin(theta[c]), cos(theta[c]), sL, sT, sigma + 6 * c);
}

// tab[2*(a - lo)] = sin(a pi/256), tab[2*(a - lo) + 1] = cos(a pi/256)
__global__ __launch_bounds__(kBlock) void k_fibre_sigma_int(int64_t n, const int *theta,
                                                            const int *phi, int lo, int hi,
                                                            const double *tab, double sL,
                                                            double sT, double *sigma,
                                                            unsigned *bad)
{
   const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
   if (c >= n) return;
   const int t = theta[c], p = phi[c];
   if (t < lo || t > hi || p < lo || p > hi) {
      atomicAdd(bad, 1u);
      return;
   }
   const double *T = tab + 2 * (t - lo), *P = tab + 2 * (p - lo);
   ibt(P[0], P[1], T[0], T[1], sL, sT, sigma + 6 * c);
}

__global__ __launch_bounds__(kBlock) void k_fgr_weights(const uint8_t *tissue, const double *sigma,
                                                        int64_t n, const int64_t *bidx,
                                                        fgr::WeightGeom G, double *A0, float *Aw,
                                                        int64_t stride, int layout)
{
   const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
   if (c >= n) return;
   float w[19];
   fgr::cell_weights(tissue, sigma, bidx[c], G, w);
   A0[c] = fgr::centre_weight(w);
#pragma unroll
   for (int k = 1; k < 19; k++)
      Aw[layout == kWeightsGrouped ? grp_index18(k - 1, c)
         : layout == kWeightsTiled ? blk_index(k - 1, c, 18)
                                   : (k - 1) * stride + c] = w[k];
}

__global__ __launch_bounds__(kBlock) void k_scatter_box(int64_t n, const uint32_t *bi, uint32_t b0,
                                                        const double *sigc, double *sigbox,
                                                        uint8_t *tissue)
{
   const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
   if (k >= n) return;
   const int64_t ib = (int64_t)(bi[k] - b0);
   tissue[ib] = 1;
#pragma unroll
   for (int q = 0; q < 6; q++) sigbox[6 * ib + q] = sigc[6 * k + q];
}

// S slab cells (padding cells share cell 0's box point, so they get cell
// 0's weights, as on the host).  Aw: the arena's tile layout; else W plain
// [18][S] and the largest |weight| (as float bits) for the compaction.
__global__ __launch_bounds__(kBlock) void k_engine_weights(const uint8_t *tissue,
                                                           const double *sigma, int64_t S,
                                                           const uint32_t *bi, uint32_t b0,
                                                           fgr::WeightGeom G, float *Aw, float *W,
                                                           unsigned *wmax)
{
   const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
   float m = 0.0f;
   if (c < S) {
      float w[19];
      fgr::cell_weights(tissue, sigma, (int64_t)(bi[c] - b0), G, w);
      if (Aw) {
#pragma unroll
         for (int k = 1; k < 19; k++) Aw[blk_index(k - 1, c, 18)] = w[k];
      } else {
#pragma unroll
         for (int k = 1; k < 19; k++) {
            W[(k - 1) * S + c] = w[k];
            m = fmaxf(m, fabsf(w[k]));
         }
      }
   }
   if (wmax) {
      for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d));
      if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(wmax, __float_as_uint(m));
   }
}

// one wavefront per (tile, neighbour pair), one cell per lane: the two
// neighbours' references (mean of the tile's nonzero tissue-cell weights,
// summed in cell order as on the host, scaled by up and rounded to FP16) and
// each cell's FP16 differences, two to a word
__global__ __launch_bounds__(kBlock) void k_compact(const float *W, int64_t S, int64_t n,
                                                    int64_t ntile, const double *upp,
                                                    float *Aref, uint32_t *Aw16)
{
   const double up = *upp;
   const int64_t g = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;   // wave-uniform
   const int lane = threadIdx.x & 63;
   if (g >= ntile * 9) return;
   const int64_t t = g / 9;
   const int p = (int)(g % 9);
   const int64_t c0 = t * kBlk, c = c0 + lane;
   const int live = (int)((n - c0) < kBlk ? (n - c0) : kBlk);   // tissue cells of the tile
   uint32_t word = 0;
   for (int h = 0; h < 2; h++) {
      const float w = W[(int64_t)(2 * p + h) * S + c];
      double sum = 0.0;
      int nz = 0;
      for (int j = 0; j < live; j++) {
         const float v = __shfl(w, j);
         if (v != 0.0f) {
            sum += v;
            ++nz;
         }
      }
      const float ref = (float)fgr::half_rne(nz ? sum / (double)nz * up : 0.0);
      if (lane == 0) Aref[t * 18 + 2 * p + h] = ref;
      const double d = (double)w * up - (double)ref;
      word |= (uint32_t)fgr::half_bits(fgr::half_rne(d)) << (16 * h);
   }
   Aw16[grp_index9(p, c)] = word;
}

}  // namespace

// the compaction's exponent from the slab's largest |weight|: 2^-wexp into
// the slab's descriptor and 2^wexp for k_compact (one thread)
__global__ void k_weight_scale(const unsigned *wmax, float *wdelta_scale, double *up)
{
   const int wexp = compact_weight_exponent(__uint_as_float(*wmax));
   *wdelta_scale = (float)ldexp(1.0, -wexp);
   *up = ldexp(1.0, wexp);
}

template <class T>
static void grow(DevBuf<T> &b, int64_t n)
{
   if ((int64_t)b.n < n) b.alloc((size_t)n);
}

void SlabSetupScratch::reserve(int64_t nbox, int64_t n, int64_t nremote, int64_t S, bool compact)
{
   grow(tissue, nbox);
   grow(sigbox, 6 * nbox);
   grow(sigc, 6 * std::max<int64_t>(n, 1));
   grow(rbi, std::max<int64_t>(nremote, 1));
   grow(rsig, 6 * std::max<int64_t>(nremote, 1));
   if (compact) {
      grow(W, 18 * S);
      grow(wmax, 1);
      grow(up, 1);
   }
}

void device_slab_weights(hipStream_t stream, SlabSetupScratch &sc, const BoxGeom &geom,
                         int64_t nbox, int64_t n, int64_t ntile, const uint32_t *bi, uint32_t b0,
                         int64_t nremote, float *Aw, float *Aref, uint32_t *Aw16,
                         float *wdelta_scale)
{
   const int64_t S = ntile * kBlk;
   CDH_REQUIRE((int64_t)sc.tissue.n >= nbox && (int64_t)sc.sigc.n >= 6 * n &&
                   (Aw || (int64_t)sc.W.n >= 18 * S),
               "device_slab_weights: scratch not reserved");
   CDH_HIP(hipMemsetAsync(sc.tissue.p, 0, nbox, stream));
   CDH_HIP(hipMemsetAsync(sc.sigbox.p, 0, sizeof(double) * 6 * nbox, stream));
   hipLaunchKernelGGL(k_scatter_box, dim3(grid_for(n)), dim3(kBlock), 0, stream, n, bi, b0,
                      sc.sigc.p, sc.sigbox.p, sc.tissue.p);
   CDH_LAUNCH_CHECK();
   if (nremote > 0) {
      hipLaunchKernelGGL(k_scatter_box, dim3(grid_for(nremote)), dim3(kBlock), 0, stream, nremote,
                         sc.rbi.p, 0u, sc.rsig.p, sc.sigbox.p, sc.tissue.p);
      CDH_LAUNCH_CHECK();
   }
   const fgr::WeightGeom G = fgr::weight_geom(geom);
   if (Aw) {
      hipLaunchKernelGGL(k_engine_weights, dim3(grid_for(S)), dim3(kBlock), 0, stream, sc.tissue.p,
                         sc.sigbox.p, S, bi, b0, G, Aw, (float *)nullptr, (unsigned *)nullptr);
      CDH_LAUNCH_CHECK();
      return;
   }
   CDH_HIP(hipMemsetAsync(sc.wmax.p, 0, sizeof(unsigned), stream));
   hipLaunchKernelGGL(k_engine_weights, dim3(grid_for(S)), dim3(kBlock), 0, stream, sc.tissue.p,
                      sc.sigbox.p, S, bi, b0, G, (float *)nullptr, sc.W.p, sc.wmax.p);
   CDH_LAUNCH_CHECK();
   hipLaunchKernelGGL(k_weight_scale, dim3(1), dim3(1), 0, stream, sc.wmax.p, wdelta_scale,
                      sc.up.p);
   CDH_LAUNCH_CHECK();
   hipLaunchKernelGGL(k_compact, dim3(grid_for(ntile * 9 * kBlk)), dim3(kBlock), 0, stream, sc.W.p,
                      S, n, ntile, sc.up.p, Aref, Aw16);
   CDH_LAUNCH_CHECK();
}

// Scratch of cdh_fibre_conductivity_int_noalloc, one copy per device for the
// life of the process: the sine/cosine table of the integer angles
// -CDH_FIBRE_ANGLE_MAX..CDH_FIBRE_ANGLE_MAX (written once, at the first call
// on the device) and, after it, the out-of-range counter (cleared by every
// call).  16,400 B of table plus one 256 B slot for the counter.
constexpr int kAngleTable = 2 * (2 * CDH_FIBRE_ANGLE_MAX + 1);
constexpr int kScratchDoubles = kAngleTable + 32;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;

__global__ void k_clear_u32(unsigned *x) { *x = 0u; }

// The device address of this device's scratch, taken once per device (with
// the device current) and the angle table written into it then.  Host code
// passes this address, never the array's name, to kernels and copies.
double *noalloc_scratch(hipStream_t s)
{
   static std::mutex mu;
   static double *addr[kMaxDevices] = {};
   int dev = -1;
   CDH_HIP(hipGetDevice(&dev));
   CDH_REQUIRE(dev >= 0 && dev < kMaxDevices, "noalloc scratch: device id out of range");
   std::lock_guard<std::mutex> lock(mu);
   if (!addr[dev]) {
      double *p = nullptr;
      CDH_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
      CDH_REQUIRE(p != nullptr, "noalloc scratch: no device address");
      static double tab[kAngleTable];
      for (int a = -CDH_FIBRE_ANGLE_MAX; a <= CDH_FIBRE_ANGLE_MAX; a++) {
         const double ang = a * (M_PI / 256);
         tab[2 * (a + CDH_FIBRE_ANGLE_MAX)] = std::sin(ang);
         tab[2 * (a + CDH_FIBRE_ANGLE_MAX) + 1] = std::cos(ang);
      }
      CDH_HIP(hipMemcpyAsync(p, tab, sizeof(tab), hipMemcpyHostToDevice, s));
      CDH_HIP(hipStreamSynchronize(s));
      addr[dev] = p;
   }
   return addr[dev];
}

void device_fgr_weights(hipStream_t stream, const uint8_t *tissue, const double *sigma,
                        int64_t ncell, const int64_t *bidx, const fgr::WeightGeom &G, double *A0,
                        float *Aw, int64_t stride, int layout)
{
   if (ncell == 0) return;
   hipLaunchKernelGGL(k_fgr_weights, dim3(grid_for(ncell)), dim3(kBlock), 0, stream, tissue,
                      sigma, ncell, bidx, G, A0, Aw, stride, layout);
   CDH_LAUNCH_CHECK();
}

}  // namespace cdh

using namespace cdh;

extern "C" {

int cdh_fibre_conductivity(int64_t n, const double *theta, const double *phi, double sigmaL,
                           double sigmaT, double *sigma, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(n >= 0 && (n == 0 || (theta && phi && sigma)),
                  "cdh_fibre_conductivity: null argument");
      if (n == 0) return;
      hipLaunchKernelGGL(k_fibre_sigma, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                         n, theta, phi, sigmaL, sigmaT, sigma);
      CDH_LAUNCH_CHECK();
   });
}

int cdh_fibre_conductivity_int(int64_t n, const int *theta, const int *phi, double sigmaL,
                               double sigmaT, double *sigma, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(n >= 0 && (n == 0 || (theta && phi && sigma)),
                  "cdh_fibre_conductivity_int: null argument");
      if (n == 0) return;
      const int lo = -CDH_FIBRE_ANGLE_MAX, hi = CDH_FIBRE_ANGLE_MAX;
      std::vector<double> tab(2 * (size_t)(hi - lo + 1));
      for (int a = lo; a <= hi; a++) {
         const double ang = a * (M_PI / 256);
         tab[2 * (a - lo)] = std::sin(ang);
         tab[2 * (a - lo) + 1] = std::cos(ang);
      }
      const hipStream_t s = (hipStream_t)stream;
      DevBuf<double> dtab(tab.size());
      DevBuf<unsigned> bad(1);
      CDH_HIP(hipMemcpyAsync(dtab.p, tab.data(), sizeof(double) * tab.size(),
                             hipMemcpyHostToDevice, s));
      CDH_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), s));
      hipLaunchKernelGGL(k_fibre_sigma_int, dim3(grid_for(n)), dim3(kBlock), 0, s, n, theta, phi,
                         lo, hi, dtab.p, sigmaL, sigmaT, sigma, bad.p);
      CDH_LAUNCH_CHECK();
      unsigned nbad = 0;
      CDH_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
      CDH_HIP(hipStreamSynchronize(s));
      CDH_REQUIRE(nbad == 0, "cdh_fibre_conductivity_int: an angle outside [-512, 512]");
   });
}

// The same without allocating: the table and the counter live in the
// library's per-device scratch.  Calls that overlap on one device race on
// the counter.
int cdh_fibre_conductivity_int_noalloc(int64_t n, const int *theta, const int *phi,
                                       double sigmaL, double sigmaT, double *sigma, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(n >= 0 && (n == 0 || (theta && phi && sigma)),
                  "cdh_fibre_conductivity_int_noalloc: null argument");
      CDH_REQUIRE(n <= INT32_MAX,
                  "cdh_fibre_conductivity_int_noalloc: more than 2^31 - 1 cells in one call");
      if (n == 0) return;
      const hipStream_t s = (hipStream_t)stream;
      double *tab = noalloc_scratch(s);
      unsigned *bad = reinterpret_cast<unsigned *>(tab + kAngleTable);
      hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(1), 0, s, bad);
      CDH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_fibre_sigma_int, dim3(grid_for(n)), dim3(kBlock), 0, s, n, theta, phi,
                         -CDH_FIBRE_ANGLE_MAX, CDH_FIBRE_ANGLE_MAX, (const double *)tab, sigmaL,
                         sigmaT, sigma, bad);
      CDH_LAUNCH_CHECK();
      unsigned nbad = 0;
      CDH_HIP(hipMemcpyAsync(&nbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s));
      CDH_HIP(hipStreamSynchronize(s));
      CDH_REQUIRE(nbad == 0, "cdh_fibre_conductivity_int_noalloc: an angle outside [-512, 512]");
   });
}

int cdh_fgr_weights(int bx, int by, int bz, double dx, double dy, double dz,
                    const uint8_t *tissue, const double *sigma, int64_t ncell,
                    const int64_t *blockIndex, double *A0, float *Aw, int64_t stride,
                    void *stream)
{
   return guard([&] {
      CDH_REQUIRE(bx >= 3 && by >= 3 && bz >= 3 && dx > 0 && dy > 0 && dz > 0,
                  "cdh_fgr_weights: bad box");
      CDH_REQUIRE(ncell >= 0 && stride >= ncell, "cdh_fgr_weights: bad sizes");
      CDH_REQUIRE(ncell == 0 || (tissue && sigma && blockIndex && A0 && Aw),
                  "cdh_fgr_weights: null argument");
      if (ncell == 0) return;
      BoxGeom g{bx, by, bz, 0, 0, 0, dx, dy, dz};
      device_fgr_weights((hipStream_t)stream, tissue, sigma, ncell, blockIndex,
                         fgr::weight_geom(g), A0, Aw, stride);
   });
}

}  // extern "C"
