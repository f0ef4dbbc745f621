// reaction.cpp -- Reaction, TT06HipReaction, reactionFactory, ReactionManager.
#include "reaction.hpp"

namespace cdh {
namespace host {

namespace {
void check(int rc)
{
   if (rc != CDH_OK) throw Error(rc, cdh_last_error());
}
[[noreturn]] void unsupported(const char *what)
{
   throw Error(kErrUnsupported, std::string("Reaction::") + what + " is not provided by this model");
}
}  // namespace

void Reaction::updateNonGate(double, const int *, int, const double *, double *)
{
   unsupported("updateNonGate");
}
void Reaction::updateGate(double, const int *, int, const double *) { unsupported("updateGate"); }
void Reaction::getCheckpointInfo(std::vector<std::string> &, std::vector<std::string> &) const
{
   unsupported("getCheckpointInfo");
}
int Reaction::getVarHandle(const std::string &) const { unsupported("getVarHandle"); }
void Reaction::setValue(int, int, double) { unsupported("setValue"); }
double Reaction::getValue(int, int) const { unsupported("getValue"); }
std::string Reaction::getUnit(const std::string &) const { unsupported("getUnit"); }

TT06HipReaction::TT06HipReaction(const cdh_tt06_params &p, int numPoints, hipStream_t stream)
    : n_(numPoints), stream_(stream)
{
   check(cdh_tt06_create(&p, numPoints, &h_));
}

TT06HipReaction::~TT06HipReaction() { cdh_tt06_destroy(h_); }

void TT06HipReaction::calc(double dt, const int *indexArray, int n, const double *Vm,
                           const double *iStim, double *dVm)
{
   check(cdh_tt06_calc(h_, dt, indexArray, n, Vm, iStim, dVm, stream_));
}

void TT06HipReaction::updateNonGate(double dt, const int *indexArray, int n, const double *Vm,
                                    double *dVR)
{
   check(cdh_tt06_update_nongate(h_, dt, indexArray, n, Vm, dVR, stream_));
}

void TT06HipReaction::updateGate(double dt, const int *indexArray, int n, const double *Vm)
{
   check(cdh_tt06_update_gate(h_, dt, indexArray, n, Vm, stream_));
}

void TT06HipReaction::initializeMembraneVoltage(const int *indexArray, int n, double *Vm)
{
   check(cdh_tt06_initialize_membrane_voltage(h_, indexArray, n, Vm, stream_));
}

void TT06HipReaction::getCheckpointInfo(std::vector<std::string> &fieldNames,
                                        std::vector<std::string> &fieldUnits) const
{
   fieldNames.clear();
   fieldUnits.clear();
   for (int i = 0; i < CDH_NSTATE; i++) {
      const char *name = nullptr, *unit = nullptr;
      int handle = 0;
      check(cdh_tt06_checkpoint_field(i, &name, &unit, &handle));
      fieldNames.push_back(name);
      fieldUnits.push_back(unit);
   }
}

int TT06HipReaction::getVarHandle(const std::string &varName) const
{
   return cdh_tt06_var_handle(varName.c_str());
}

void TT06HipReaction::setValue(int iCell, int varHandle, double value)
{
   CDH_HIP(hipStreamSynchronize(stream_));
   check(cdh_tt06_set_value(h_, iCell, varHandle, value));
}

double TT06HipReaction::getValue(int iCell, int varHandle) const
{
   CDH_HIP(hipStreamSynchronize(stream_));
   double v = 0.0;
   check(cdh_tt06_get_value(h_, iCell, varHandle, &v));
   return v;
}

std::string TT06HipReaction::getUnit(const std::string &varName) const
{
   for (int i = 0; i < CDH_NSTATE; i++) {
      const char *name = nullptr, *unit = nullptr;
      int handle = 0;
      check(cdh_tt06_checkpoint_field(i, &name, &unit, &handle));
      if (varName == name) return unit;
   }
   return "1";
}

void TT06HipReaction::getState(double *host) const
{
   CDH_HIP(hipStreamSynchronize(stream_));
   check(cdh_tt06_get_state(h_, host));
}

void TT06HipReaction::setState(const double *host)
{
   CDH_HIP(hipStreamSynchronize(stream_));
   check(cdh_tt06_set_state(h_, host));
}

std::unique_ptr<Reaction> reactionFactory(const std::string &method, const cdh_tt06_params &p,
                                          int numPoints, hipStream_t stream)
{
   if (method == "TT06Dev" || method == "TT06Hip")
      return std::make_unique<TT06HipReaction>(p, numPoints, stream);
   throw Error(kErrArg, "reactionFactory: unknown reaction method " + method);
}

void ReactionManager::addReaction(std::unique_ptr<Reaction> r, const std::vector<int> &indices)
{
   auto g = std::make_unique<Group>();
   g->r = std::move(r);
   g->host_idx = indices;
   g->idx.alloc(indices.empty() ? 1 : indices.size());
   g->idx.upload(indices.data(), indices.size());
   groups_.push_back(std::move(g));
}

void ReactionManager::calc(double dt, const double *Vm, const double *iStim, double *dVm)
{
   for (auto &g : groups_) g->r->calc(dt, g->idx.p, (int)g->host_idx.size(), Vm, iStim, dVm);
}

void ReactionManager::initializeMembraneVoltage(double *Vm)
{
   for (auto &g : groups_)
      g->r->initializeMembraneVoltage(g->idx.p, (int)g->host_idx.size(), Vm);
}

}  // namespace host
}  // namespace cdh
