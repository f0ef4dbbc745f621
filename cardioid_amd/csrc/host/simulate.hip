// simulate.hip -- BoxStimulus, ActivationSensor and simulationLoop.
#include "simulate.hpp"

#include <cmath>

namespace cdh {
namespace host {

namespace {
void check(int rc)
{
   if (rc != CDH_OK) throw Error(rc, cdh_last_error());
}
}  // namespace

BoxStimulus::BoxStimulus(const cdh_box_stim &b, const std::vector<int> &cells, hipStream_t stream)
    : b_(b), n_((int)cells.size()), stream_(stream)
{
   list_.alloc(cells.empty() ? 1 : cells.size());
   list_.upload(cells.data(), cells.size());
}

void BoxStimulus::stim(double time, double *iStim)
{
   if (!(time > b_.t0 && time < b_.tf)) return;
   const double t = time - std::floor(time / b_.period) * b_.period;
   if (!(t >= b_.phase && t < b_.phase + b_.duration)) return;
   if (b_.amplitude == 0.0 || n_ == 0) return;
   // subClassStim assigns the value on its cells (BoxStimulus.cc:61-62)
   check(cdh_set_list(n_, list_.p, b_.amplitude, iStim, stream_));
}

ActivationSensor::ActivationSensor(int64_t n, double threshold, const double *vrestForApd,
                                   hipStream_t stream)
    : stream_(stream)
{
   check(cdh_activation_sensor_create(n, threshold, &h_));
   if (vrestForApd) check(cdh_activation_sensor_enable_apd(h_, vrestForApd));
}

ActivationSensor::~ActivationSensor() { cdh_activation_sensor_destroy(h_); }

void ActivationSensor::eval(double time, const double *Vm)
{
   check(cdh_activation_sensor_eval(h_, time, Vm, stream_));
}

void ActivationSensor::get(double *tact0, double *tact, double *trec) const
{
   CDH_HIP(hipStreamSynchronize(stream_));
   check(cdh_activation_sensor_get(h_, tact0, tact, trec));
}

void ActivationSensor::counts(int64_t *n_tact0, int64_t *n_tact, int64_t *n_trec) const
{
   CDH_HIP(hipStreamSynchronize(stream_));
   check(cdh_activation_sensor_counts(h_, n_tact0, n_tact, n_trec));
}

void ActivationSensor::t90(double *out) const
{
   CDH_HIP(hipStreamSynchronize(stream_));
   check(cdh_activation_sensor_get_t90(h_, out));
}

void simulationLoop(Simulate &sim, int nsteps)
{
   const int64_t n = sim.nLocal;
   for (int step = 0; step < nsteps; step++) {
      // iStim = 0, then every stimulus (simulationLoop.cc:186-193)
      check(cdh_fill(n, 0.0, sim.iStim.p, sim.stream));
      for (auto &s : sim.stimulus) s->stim(sim.time, sim.iStim.p);
      // reaction (:195-200)
      sim.reaction->calc(sim.dt, sim.Vm.p, sim.iStim.p, sim.dVmReaction.p);
      // diffusion (:203-210); one rank, so no remote cells arrive
      sim.diffusion->updateLocalVoltage(sim.Vm.p);
      sim.diffusion->updateRemoteVoltage(nullptr);
      sim.diffusion->calc(sim.dVmDiffusion.p);
      // range check before the integrator (:214-217)
      if (sim.checkRange && n > 0) {
         hipLaunchKernelGGL(k_check_range_f64, dim3(grid_for(n, 256)), dim3(256), 0, sim.stream,
                            (int)n, sim.Vm.p, sim.dVmReaction.p, sim.dVmDiffusion.p, sim.vMin,
                            sim.vMax, sim.loop, sim.range.p);
         CDH_LAUNCH_CHECK();
      }
      // integrator (:218-227), then time and loop (:229-231)
      check(cdh_integrate(n, sim.dt, sim.Vm.p, sim.dVmReaction.p, sim.dVmDiffusion.p,
                          sim.iStim.p, sim.stream));
      sim.time += sim.dt;
      ++sim.loop;
      // sensors at the new time (:233-241)
      for (auto &s : sim.sensor) s->eval(sim.time, sim.Vm.p);
   }
}

}  // namespace host
}  // namespace cdh
