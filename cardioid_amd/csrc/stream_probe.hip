// stream_probe.hip -- measured HBM stream rates of the fused FP64 kernel's
// state pattern, for the roofline's practical ceiling (cdh_stream_rate).
//
// The pattern: 19 FP64 fields per cell in runs of kRun cells with the
// fields in pairs (run_pair_index), one wavefront per 64 cells, one
// non-temporal 16-byte access per lane and pair (8 bytes for the odd
// field), as k_engine_f64 streams its states.  Read only, write only, or
// read and written in place.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cardioid_hip.h"
#include "box_access.hpp"
#include "hip_util.hpp"
#include "tt06_model.hpp"

namespace cdh {
namespace {

constexpr int kFields = N_STATE;

template <int PATTERN>   // 0 read, 1 write, 2 read and write in place
__global__ __launch_bounds__(256) void k_stream(double *__restrict__ s, int64_t ntile,
                                                double *__restrict__ sink)
{
   const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
   if (t >= ntile) return;
   const int64_t k0 = t * kBlk, r0 = k0 & ~(kRun - 1);
   const uint32_t rw = (uint32_t)run_width(k0, ntile * kBlk);
   const __amdgpu_buffer_rsrc_t R = stream_rsrc(s + r0 * kFields);
   const uint32_t lc = (uint32_t)(k0 - r0 + (threadIdx.x & 63));
   double v[kFields];
   if (PATTERN != 1) {
#pragma unroll
      for (int q = 0; q + 1 < kFields; q += 2) stream_load_f64x2(R, 16u * lc, 8u * rw * q, v[q], v[q + 1]);
      v[kFields - 1] = stream_load_f64(R, 8u * lc, 8u * rw * (kFields - 1));
   } else {
#pragma unroll
      for (int q = 0; q < kFields; q++) v[q] = (double)(t + q);
   }
   double acc = 0.0;
#pragma unroll
   for (int q = 0; q < kFields; q++) acc += v[q];
   if (PATTERN == 0) {
      // keeps the loads; the buffer holds zeros, so the sum never matches
      if (acc == 1.0) sink[0] = acc;
      return;
   }
#pragma unroll
   for (int q = 0; q + 1 < kFields; q += 2)
      stream_store_f64x2(v[q] + acc * 1e-300, v[q + 1] + acc * 1e-300, R, 16u * lc, 8u * rw * q);
   stream_store_f64(v[kFields - 1] + acc * 1e-300, R, 8u * lc, 8u * rw * (kFields - 1));
}

}  // namespace
}  // namespace cdh

using namespace cdh;

extern "C" int cdh_stream_rate(int pattern, int64_t cells, int reps, double *gbps)
{
   return guard([&] {
      CDH_REQUIRE(pattern >= 0 && pattern <= 2, "cdh_stream_rate: pattern is 0, 1 or 2");
      CDH_REQUIRE(cells > 0 && reps > 0 && gbps, "cdh_stream_rate: bad arguments");
      const int64_t ntile = (cells + kBlk - 1) / kBlk;
      const size_t bytes = (size_t)ntile * kFields * kBlk * sizeof(double);
      DevBuf<double> buf(bytes / sizeof(double));
      DevBuf<double> sink(1);
      hipStream_t st = nullptr;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      CDH_HIP(hipStreamCreate(&st));
      try {
         CDH_HIP(hipMemsetAsync(buf.p, 0, bytes, st));
         CDH_HIP(hipEventCreate(&e0));
         CDH_HIP(hipEventCreate(&e1));
         const int64_t nblock = (ntile + 3) / 4;
         CDH_REQUIRE(nblock <= 0x7fffffff, "cdh_stream_rate: too many cells");
         auto launch = [&] {
            const dim3 g((unsigned)nblock), b(256);
            if (pattern == 0)
               hipLaunchKernelGGL(k_stream<0>, g, b, 0, st, buf.p, ntile, sink.p);
            else if (pattern == 1)
               hipLaunchKernelGGL(k_stream<1>, g, b, 0, st, buf.p, ntile, sink.p);
            else
               hipLaunchKernelGGL(k_stream<2>, g, b, 0, st, buf.p, ntile, sink.p);
            CDH_HIP(hipGetLastError());
         };
         for (int r = 0; r < 3; r++) launch();
         CDH_HIP(hipEventRecord(e0, st));
         for (int r = 0; r < reps; r++) launch();
         CDH_HIP(hipEventRecord(e1, st));
         CDH_HIP(hipEventSynchronize(e1));
         float ms = 0.0f;
         CDH_HIP(hipEventElapsedTime(&ms, e0, e1));
         const double moved = (double)bytes * (pattern == 2 ? 2.0 : 1.0) * reps;
         *gbps = moved / (ms * 1e-3) / 1e9;
      } catch (...) {
         if (e0) (void)hipEventDestroy(e0);
         if (e1) (void)hipEventDestroy(e1);
         (void)hipStreamSynchronize(st);
         (void)hipStreamDestroy(st);
         throw;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipStreamDestroy(st);
   });
}
