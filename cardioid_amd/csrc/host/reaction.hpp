// reaction.hpp -- the reference's Reaction interface and ReactionManager
// over device arrays.
//
// Reaction mirrors elec/Reaction.hh:10-43: the same methods with the same
// meaning, the lazy_array arguments replaced by HIP device pointers and a
// count (what lazy_array::useOn(GPU) hands out).  Where the reference's base
// class asserts (elec/Reaction.cc:61-117), these throw an Error with
// kErrUnsupported.  ReactionManager mirrors elec/ReactionManager.cc: one
// reaction object per cell type, each with the index array of its cells in
// ascending local order (:184-261), calc and initializeMembraneVoltage
// fanned out over them (:15-65).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../../include/cardioid_hip.h"
#include "../hip_util.hpp"

namespace cdh {
namespace host {

class Reaction {
 public:
   virtual ~Reaction() = default;
   virtual std::string methodName() const = 0;
   // reads Vm[indexArray[i]] and writes dVm[indexArray[i]] for i < n
   virtual void calc(double dt, const int *indexArray, int n, const double *Vm,
                     const double *iStim, double *dVm) = 0;
   virtual void updateNonGate(double dt, const int *indexArray, int n, const double *Vm,
                              double *dVR);
   virtual void updateGate(double dt, const int *indexArray, int n, const double *Vm);
   virtual void initializeMembraneVoltage(const int *indexArray, int n, double *Vm) = 0;
   virtual void getCheckpointInfo(std::vector<std::string> &fieldNames,
                                  std::vector<std::string> &fieldUnits) const;
   virtual int getVarHandle(const std::string &varName) const;
   virtual void setValue(int iCell, int varHandle, double value);
   virtual double getValue(int iCell, int varHandle) const;
   virtual std::string getUnit(const std::string &varName) const;
};

// TT06Dev (configuration B) on the device through cdh_tt06_*
class TT06HipReaction : public Reaction {
 public:
   TT06HipReaction(const cdh_tt06_params &p, int numPoints, hipStream_t stream);
   ~TT06HipReaction() override;
   std::string methodName() const override { return "TT06Dev"; }
   void calc(double dt, const int *indexArray, int n, const double *Vm, const double *iStim,
             double *dVm) override;
   void updateNonGate(double dt, const int *indexArray, int n, const double *Vm,
                      double *dVR) override;
   void updateGate(double dt, const int *indexArray, int n, const double *Vm) override;
   void initializeMembraneVoltage(const int *indexArray, int n, double *Vm) override;
   void getCheckpointInfo(std::vector<std::string> &fieldNames,
                          std::vector<std::string> &fieldUnits) const override;
   int getVarHandle(const std::string &varName) const override;
   void setValue(int iCell, int varHandle, double value) override;
   double getValue(int iCell, int varHandle) const override;
   std::string getUnit(const std::string &varName) const override;
   // bulk state transfer, host [19][numPoints]
   void getState(double *host) const;
   void setState(const double *host);
   int numPoints() const { return n_; }

 private:
   cdh_tt06 *h_ = nullptr;
   int n_ = 0;
   hipStream_t stream_ = nullptr;
};

// reactionFactory (elec/reactionFactory.cc:30-80): the method name selects
// the model; unknown names are an error (the reference asserts)
std::unique_ptr<Reaction> reactionFactory(const std::string &method, const cdh_tt06_params &p,
                                          int numPoints, hipStream_t stream);

class ReactionManager {
 public:
   ReactionManager() = default;
   // the reaction serving the local cells `indices` (ascending)
   void addReaction(std::unique_ptr<Reaction> r, const std::vector<int> &indices);
   void calc(double dt, const double *Vm, const double *iStim, double *dVm);
   void initializeMembraneVoltage(double *Vm);
   int size() const { return (int)groups_.size(); }
   Reaction &reaction(int k) { return *groups_[k]->r; }
   const std::vector<int> &indices(int k) const { return groups_[k]->host_idx; }

 private:
   struct Group {
      std::unique_ptr<Reaction> r;
      std::vector<int> host_idx;
      DevBuf<int> idx;
   };
   std::vector<std::unique_ptr<Group>> groups_;
};

}  // namespace host
}  // namespace cdh
