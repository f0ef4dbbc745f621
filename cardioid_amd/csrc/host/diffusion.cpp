// diffusion.cpp -- HipFGRDiffusion and diffusionFactory.
#include "diffusion.hpp"

namespace cdh {
namespace host {

namespace {
void check(int rc)
{
   if (rc != CDH_OK) throw Error(rc, cdh_last_error());
}
}  // namespace

HipFGRDiffusion::HipFGRDiffusion(const cdh_fgr_params &p, hipStream_t stream)
    : Diffusion(p.diffusionScale), stream_(stream), nRemote_(p.nRemote)
{
   check(cdh_fgr_create(&p, &h_));
}

HipFGRDiffusion::~HipFGRDiffusion() { cdh_fgr_destroy(h_); }

void HipFGRDiffusion::updateLocalVoltage(const double *VmLocal)
{
   check(cdh_fgr_update_local_voltage(h_, VmLocal, stream_));
}

void HipFGRDiffusion::updateRemoteVoltage(const double *VmRemote)
{
   if (nRemote_ > 0) check(cdh_fgr_update_remote_voltage(h_, VmRemote, stream_));
}

void HipFGRDiffusion::calc(double *dVm) { check(cdh_fgr_calc(h_, dVm, stream_)); }

std::unique_ptr<Diffusion> diffusionFactory(const std::string &method, const cdh_fgr_params &p,
                                            hipStream_t stream)
{
   if (method == "FGR" || method == "fgr" || method == "hip")
      return std::make_unique<HipFGRDiffusion>(p, stream);
   throw Error(kErrArg, "diffusionFactory: unknown diffusion method " + method);
}

}  // namespace host
}  // namespace cdh
