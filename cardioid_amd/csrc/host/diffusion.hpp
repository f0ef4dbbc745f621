// diffusion.hpp -- the reference's Diffusion interface over device arrays,
// with FGR on the device through cdh_fgr_*.
//
// Diffusion mirrors elec/Diffusion.hh:19-39 (omp-loop contract: calc
// assigns dVm); diffusionFactory mirrors the method switch of
// elec/diffusionFactory.cc:38-68, where "FGR" and "hip" select this
// implementation and any other name is an error (the reference asserts).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "../../../include/cardioid_hip.h"
#include "../hip_util.hpp"

namespace cdh {
namespace host {

class Diffusion {
 public:
   explicit Diffusion(double diffusionScale) : scale_(diffusionScale) {}
   virtual ~Diffusion() = default;
   virtual void updateLocalVoltage(const double *VmLocal) = 0;
   virtual void updateRemoteVoltage(const double *VmRemote) = 0;
   virtual void calc(double *dVm) = 0;
   double diffusionScale() const { return scale_; }

 private:
   double scale_;
};

class HipFGRDiffusion : public Diffusion {
 public:
   HipFGRDiffusion(const cdh_fgr_params &p, hipStream_t stream);
   ~HipFGRDiffusion() override;
   void updateLocalVoltage(const double *VmLocal) override;
   void updateRemoteVoltage(const double *VmRemote) override;
   void calc(double *dVm) override;

 private:
   cdh_fgr *h_ = nullptr;
   hipStream_t stream_ = nullptr;
   int64_t nRemote_ = 0;
};

std::unique_ptr<Diffusion> diffusionFactory(const std::string &method, const cdh_fgr_params &p,
                                            hipStream_t stream);

}  // namespace host
}  // namespace cdh
