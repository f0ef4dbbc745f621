// simulate.hpp -- Simulate and simulationLoop (omp loop) over the plugin
// objects, as the reference structures them.
//
// Simulate holds what elec/Simulate.hh holds for this path: dt, time, loop,
// the local Vm / iStim / dVmReaction / dVmDiffusion arrays (device, local
// cell order), the ReactionManager, the Diffusion, the stimuli, the
// sensors and the range check.  simulationLoop advances it by the steps of
// elec/simulationLoop.cc:170-241.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../hip_util.hpp"
#include "../kernels.hpp"
#include "diffusion.hpp"
#include "reaction.hpp"

namespace cdh {
namespace host {

// Stimulus::stim (elec/Stimulus.hh:21-27) with a BoxStimulus cell list
// (elec/BoxStimulus.cc:18-64) and a PeriodicPulse (elec/PeriodicPulse.cc:22-28)
class BoxStimulus {
 public:
   BoxStimulus(const cdh_box_stim &b, const std::vector<int> &cells, hipStream_t stream);
   void stim(double time, double *iStim);

 private:
   cdh_box_stim b_;
   int n_ = 0;
   DevBuf<int> list_;
   hipStream_t stream_;
};

// ActivationTimeSensor and ActivationAndRecoverySensor
// (elec/ActivationTimeSensor.cc:99-111, elec/ActivationAndRecoverySensor.cc:
// 76-92) evaluated every step, on the device through cdh_activation_sensor_*
class ActivationSensor {
 public:
   ActivationSensor(int64_t n, double threshold, const double *vrestForApd, hipStream_t stream);
   ~ActivationSensor();
   void eval(double time, const double *Vm);
   void get(double *tact0, double *tact, double *trec) const;
   void counts(int64_t *n_tact0, int64_t *n_tact, int64_t *n_trec) const;
   void t90(double *out) const;

 private:
   cdh_sensor *h_ = nullptr;
   hipStream_t stream_;
};

struct Simulate {
   double dt = 0.0, time = 0.0;
   int loop = 0;
   int64_t nLocal = 0;
   hipStream_t stream = nullptr;
   DevBuf<double> Vm, iStim, dVmReaction, dVmDiffusion;
   std::unique_ptr<ReactionManager> reaction;
   std::unique_ptr<Diffusion> diffusion;
   std::vector<std::unique_ptr<BoxStimulus>> stimulus;
   std::vector<std::unique_ptr<ActivationSensor>> sensor;
   // Simulate::checkRanges (elec/Simulate.cc:35-53, keywords checkRanges,
   // VmMin, VmMax of elec/initializeSimulate.cc:96-98)
   int checkRange = 1;
   double vMin = -150.0, vMax = 100.0;
   DevBuf<RangeRec> range;

   explicit Simulate(hipStream_t s) : stream(s) {}
};

// nsteps omp-loop steps (elec/simulationLoop.cc:170-241); asynchronous on
// the simulation's stream
void simulationLoop(Simulate &sim, int nsteps);

}  // namespace host
}  // namespace cdh
